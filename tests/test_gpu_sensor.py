"""GPU parity of the sensor path (b3amd_test_overlap, b3amd_mesh_contacts_test_overlap)
against the oracle: b3TestOverlap (collide.cpp:84-95) with each contact's simplex cache
carried across steps (gjk.cpp:743-841), bit-exact in the overlap flags, the cache records
and the b3_gjk* statistics counters including b3_gjkCacheHits; mixed with ordinary contacts
as b3Contact::Update mixes them (contact.cpp:55-61)."""
import sys

import numpy as np
import pytest

import bounce_amd
from bounce_amd.types import empty_caches
from util import (assert_caches_equal, assert_manifolds_equal, defined_hull_cache, defined_simplex_cache,
                  mixed_workload)

pytestmark = pytest.mark.gpu

SIMPLEX_FIELDS = ("metric", "iterations", "count", "sindex1", "sindex2")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def moved(xfs, rng, step):
    from test_sensor_oracle import moved as mv
    return mv(xfs, rng, step)


def test_gpu_test_overlap_equals_oracle(ctx, lib, oracle):
    """Four moving steps of 4,000 sphere/capsule/hull pairs (hull-hull included) with caches
    carried: flags, caches and counters equal the oracle's."""
    say("case: sensor overlap, 4 steps with simplex caches")
    pool, shapes, pairs, xfs = mixed_workload(lib, 711, 4000, gap_mean=0.0, gap_sigma=0.05)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    rng = np.random.default_rng(4)
    cg, co = empty_caches(len(pairs)), empty_caches(len(pairs))
    ctx.set_counting(True)
    try:
        ctx.reset_counters()
        oracle.reset_counters()
        x = xfs
        for step in range(4):
            x = moved(x, rng, step)
            got = ctx.test_overlap(pairs, x, caches=cg)
            want = oracle.test_overlap(pool, shapes, pairs, x, caches=co)
            np.testing.assert_array_equal(got, want, "step %d" % step)
            assert cg.tobytes() == co.tobytes(), "step %d: caches differ" % step
        got_c, want_c = ctx.counters(), oracle.counters()
    finally:
        ctx.set_counting(False)
    for k in ("gjk_calls", "gjk_iters", "gjk_max_iters", "gjk_cache_hits"):
        assert got_c[k] == want_c[k], (k, got_c[k], want_c[k])
    assert want_c["gjk_cache_hits"] > len(pairs)
    fresh = ctx.test_overlap(pairs, x)  # no caches: the empty-cache simplex
    np.testing.assert_array_equal(fresh, oracle.test_overlap(pool, shapes, pairs, x))


def test_gpu_test_overlap_special_cases(ctx, lib, oracle):
    """Coincident origins, deep overlaps and jumps that flush the cached simplex, and invalid
    pairs (a shape index outside the table) flagged rather than read."""
    say("case: sensor overlap special cases")
    pool, shapes, pairs, xfs = mixed_workload(lib, 712, 800, gap_mean=-0.05, gap_sigma=0.2)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    x = xfs.copy()
    x[::7, 1, 9:12] = x[::7, 0, 9:12]
    cg, co = empty_caches(len(pairs)), empty_caches(len(pairs))
    for step, scale in enumerate((1.0, 1.0, 3.0, 0.3)):
        xs = x.copy()
        xs[:, 1, 9:12] = xs[:, 0, 9:12] + scale * (xs[:, 1, 9:12] - xs[:, 0, 9:12])
        got = ctx.test_overlap(pairs, xs, caches=cg)
        want = oracle.test_overlap(pool, shapes, pairs, xs, caches=co)
        np.testing.assert_array_equal(got, want, "step %d" % step)
        assert cg.tobytes() == co.tobytes(), "step %d" % step
    bad = pairs[:4].copy()
    bad[1, 1] = len(shapes) + 5
    with pytest.raises(bounce_amd.BounceError):
        ctx.test_overlap(bad, x[:4])  # the host entry point validates pairs


def test_gpu_sensor_and_ordinary_contacts_mixed(ctx, lib, oracle):
    """A world's contacts where every third one has a sensor shape: those run b3TestOverlap,
    the others the narrowphase, each with its own cache, over three moving steps with sweeps."""
    say("case: sensor and ordinary contacts, 3 steps")
    pool, shapes, pairs, xfs = mixed_workload(lib, 713, 3000, gap_mean=-0.01, gap_sigma=0.05)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    sensor = np.arange(len(pairs)) % 3 == 0
    ps, po = pairs[sensor], pairs[~sensor]
    rng = np.random.default_rng(8)
    cs_g, cs_o = empty_caches(len(ps)), empty_caches(len(ps))
    cc_g, cc_o = empty_caches(len(po)), empty_caches(len(po))
    x = xfs
    for step in range(3):
        x = moved(x, rng, step)
        sw = np.zeros((len(po), 4, 4), np.float32)
        sw[:, :, 3] = 1.0
        tilt = rng.uniform(0, 0.04, len(po)).astype(np.float32)
        sw[:, 3, 2], sw[:, 3, 3] = tilt, np.sqrt(1 - tilt * tilt)
        np.testing.assert_array_equal(ctx.test_overlap(ps, x[sensor], caches=cs_g),
                                      oracle.test_overlap(pool, shapes, ps, x[sensor], caches=cs_o), "step %d" % step)
        assert cs_g.tobytes() == cs_o.tobytes()
        got = ctx.collide(po, x[~sensor], sweeps=sw, caches=cc_g)
        want = oracle.collide(pool, shapes, po, x[~sensor], sweeps=sw, caches=cc_o)
        assert_manifolds_equal(got, want, "step %d" % step)
        assert_caches_equal(defined_hull_cache(cc_g), defined_hull_cache(cc_o), "step %d" % step)


def test_gpu_test_overlap_device_api(ctx, lib):
    say("case: sensor overlap, device-resident API")
    pool, shapes, pairs, xfs = mixed_workload(lib, 714, 2000)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    caches = empty_caches(len(pairs))
    want = ctx.test_overlap(pairs, xfs, caches=caches.copy())
    n = len(pairs)
    dp, dx, dc, do = ctx.alloc(pairs.nbytes), ctx.alloc(xfs.nbytes), ctx.alloc(caches.nbytes), ctx.alloc(4 * n)
    try:
        dp.upload(pairs)
        dx.upload(xfs)
        dc.upload(caches)
        import ctypes
        f = lib.b3amd_test_overlap_device
        f.argtypes = [ctypes.c_void_p, ctypes.c_uint32] + [ctypes.c_void_p] * 4
        assert f(ctx.handle, n, dp.ptr, dx.ptr, dc.ptr, do.ptr) == 0
        ctx.synchronize()
        np.testing.assert_array_equal(do.download(np.zeros(n, np.uint32)), want)
    finally:
        for b in (dp, dx, dc, do):
            b.free()


def test_gpu_mesh_test_overlap_equals_oracle(ctx, lib, oracle):
    """Sensor mesh contacts over a 48 x 48 heightfield, six moving steps with FindNewPairs in
    between: flags, every triangle cache (defined fields) and the counters equal the oracle's
    b3MeshContact::TestOverlap."""
    import oracle_lib
    from util import mesh_workload, moving_mesh_steps
    say("case: sensor mesh contacts, 6 steps")
    pool, shapes, v, t, xfs = mesh_workload(lib, 93, 48, 48, 300)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    mid = ctx.upload_mesh(v, t)
    sa = np.arange(len(shapes), dtype=np.uint32)
    steps = moving_mesh_steps(xfs, 7, 6)
    ref = oracle_lib.RefMeshContacts(oracle.lib, pool, shapes, v, t, sa, steps[0][0], prefix="b3o_")
    ctx.set_counting(True)
    try:
        ctx.reset_counters()
        oracle.reset_counters()
        with bounce_amd.MeshContacts(ctx, mid, sa, steps[0][0]) as dev:
            for s in range(6):
                x, sw, p = steps[s]
                np.testing.assert_array_equal(dev.test_overlap(x), ref.test_overlap(x), "step %d" % s)
                dc, dl, dca, _ = dev.read()
                rc, rl, rca, _ = ref.read()
                np.testing.assert_array_equal(dc, rc)
                for i in range(len(dc)):
                    k = min(int(dc[i]), 64)
                    assert_caches_equal(defined_simplex_cache(dca[i, :k]), defined_simplex_cache(rca[i, :k]),
                                        "step %d contact %d" % (s, i), fields=SIMPLEX_FIELDS)
                dev.synchronize(steps[s + 1][0], p)
                ref.synchronize(steps[s + 1][0], p)
        got_c, want_c = ctx.counters(), oracle.counters()
    finally:
        ctx.set_counting(False)
        ref.close()
    for k in ("gjk_calls", "gjk_iters", "gjk_max_iters", "gjk_cache_hits"):
        assert got_c[k] == want_c[k], (k, got_c[k], want_c[k])
    assert want_c["gjk_cache_hits"] > 0
