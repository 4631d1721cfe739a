"""The hull registry (b3amd_hull_add / b3amd_hull_free: handles in place of the b3Hull
pointer a b3HullShape borrows, hull_shape.h:44), the single-pair b3CollideShapeAndShape
signature (collide.h:63-66) and the contact-cache arrays (convex_contact.cpp:35-36), each
checked against the oracle or against the batch path, bit-exact."""
import sys

import numpy as np
import pytest

import bounce_amd
from bounce_amd.types import CACHE_DTYPE, HullPool, empty_caches
from util import assert_caches_equal, assert_manifolds_equal, defined_hull_cache, hull_workload, mixed_workload

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def subset(pool, idx):
    return HullPool.concat([pool.hull(int(i)) for i in idx])


def test_gpu_registry_add_free_reuse(ctx, lib, oracle):
    """Hulls registered in pieces, some freed and others added into the freed slots in a
    shuffled order: a batch over the handles equals the oracle over the source pool."""
    say("case: hull registry add/free/reuse, 4,000 pairs")
    pool, shapes, pairs, xfs = hull_workload(lib, 811, 64, 4000)
    ctx.upload_hulls(subset(pool, range(32)))
    info0 = ctx.hull_registry()
    assert info0["slots"] == 32 and info0["live"] == 32 and info0["capacity_bytes"] == info0["used_bytes"]
    h = ctx.add_hulls(subset(pool, range(32, 48)))
    np.testing.assert_array_equal(h, np.arange(32, 48))
    info1 = ctx.hull_registry()
    assert info1["capacity_bytes"] >= 2 * info0["capacity_bytes"] and info1["capacity_bytes"] >= info1["used_bytes"]
    # A hull in use by a shape cannot be freed; a freed hull cannot be used.
    s = np.zeros(1, shapes.dtype)
    s["type"], s["hull"] = bounce_amd.types.HULL, 5
    ctx.upload_shapes(s)
    with pytest.raises(bounce_amd.BounceError):
        ctx.free_hull(5)
    ctx.upload_shapes(s[:0])
    for k in (5, 9, 40):
        ctx.free_hull(k)
    with pytest.raises(bounce_amd.BounceError):
        ctx.free_hull(9)
    with pytest.raises(bounce_amd.BounceError):
        ctx.upload_shapes(s)  # hull 5 was freed
    info2 = ctx.hull_registry()
    assert info2["slots"] == 48 and info2["live"] == 45 and info2["used_bytes"] < info1["used_bytes"]
    # Re-add the three in another order (handles 5, 9, 40 are reused, lowest first), then the rest.
    src = [40, 5, 9] + list(range(48, 64))
    h = ctx.add_hulls(subset(pool, src))
    np.testing.assert_array_equal(h, [5, 9, 40] + list(range(48, 64)))
    source_of = np.arange(64)
    source_of[h] = src
    handle_of = np.argsort(source_of)
    info3 = ctx.hull_registry()
    assert info3["slots"] == 64 and info3["live"] == 64
    sh = pool.shapes()  # shape i is handle i
    ctx.upload_shapes(sh)
    gp = handle_of[pairs].astype(np.uint32)  # source hull -> handle (shape i is handle i)
    cg, co = empty_caches(len(pairs)), empty_caches(len(pairs))
    got = ctx.collide(gp, xfs, caches=cg)
    want = oracle.collide(pool, shapes, pairs, xfs, caches=co)
    assert_manifolds_equal(got, want, "registry batch")
    assert_caches_equal(defined_hull_cache(cg), defined_hull_cache(co), "registry caches")
    assert int(want["pointCount"].sum()) > 1000


def test_gpu_registry_growth_keeps_records(ctx, lib, oracle):
    """Adding hulls one at a time grows the arena by doubling; every record survives the moves."""
    say("case: hull registry growth, 48 single adds")
    pool, shapes, pairs, xfs = hull_workload(lib, 812, 48, 2000)
    ctx.upload_hulls(subset(pool, [0]))
    caps = set()
    for i in range(1, 48):
        assert ctx.add_hulls(subset(pool, [i]))[0] == i
        caps.add(ctx.hull_registry()["capacity_bytes"])
    assert len(caps) >= 4
    ctx.upload_shapes(shapes)
    np.testing.assert_array_equal(ctx.collide(pairs, xfs).view(np.uint8), oracle.collide(pool, shapes, pairs, xfs).view(np.uint8))


def test_gpu_collide_shape_and_shape_equals_batch(ctx, lib):
    """b3amd_collide_shape_and_shape, one pair per call with its cache and sweeps, over two
    steps: every manifold and cache equals the batch's."""
    say("case: single-pair entry point, 400 mixed pairs, 2 steps")
    pool, shapes, pairs, xfs = mixed_workload(lib, 813, 400, gap_mean=-0.02, gap_sigma=0.05)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    rng = np.random.default_rng(3)
    cb, cs = empty_caches(len(pairs)), empty_caches(len(pairs))
    for step in range(2):
        sw = np.zeros((len(pairs), 4, 4), np.float32)
        sw[:, :, 3] = 1.0
        tilt = rng.uniform(0, 0.04, len(pairs)).astype(np.float32)
        sw[:, 3, 2], sw[:, 3, 3] = tilt, np.sqrt(1 - tilt * tilt)
        batch = ctx.collide(pairs, xfs, sweeps=sw, caches=cb)
        one = np.zeros(len(pairs), batch.dtype)
        for i in range(len(pairs)):
            one[i] = ctx.collide_one(xfs[i, 0], pairs[i, 0], xfs[i, 1], pairs[i, 1], sweeps=sw[i], cache=cs[i:i + 1])
        assert one.tobytes() == batch.tobytes(), "step %d" % step
        assert cs.tobytes() == cb.tobytes(), "step %d" % step
    with pytest.raises(bounce_amd.BounceError):
        ctx.collide_one(xfs[0, 0], len(shapes) + 3, xfs[0, 1], pairs[0, 1])


def test_gpu_cache_alloc_reset(ctx, lib):
    """Caches from b3amd_cache_alloc start as a new contact's; a batch fills them; a reset
    empties exactly its range."""
    say("case: cache alloc/reset/free")
    pool, shapes, pairs, xfs = hull_workload(lib, 814, 32, 3000, gap_mean=-0.02)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    n = len(pairs)
    ca = ctx.cache_alloc(n)
    dp, dx, do = ctx.alloc(pairs.nbytes), ctx.alloc(xfs.nbytes), ctx.alloc(240 * n)
    try:
        got = ca.download(np.zeros(n, CACHE_DTYPE))
        assert got.tobytes() == empty_caches(n).tobytes()
        dp.upload(pairs)
        dx.upload(xfs)
        ctx.collide_device(n, dp, dx, do, d_caches=ca)
        ctx.synchronize()
        host = empty_caches(n)
        ctx.collide(pairs, xfs, caches=host)
        filled = ca.download(np.zeros(n, CACHE_DTYPE))
        assert filled.tobytes() == host.tobytes()
        assert (filled["state"] != 2).sum() > n // 2
        ca.reset(100, 500)
        ctx.synchronize()
        after = ca.download(np.zeros(n, CACHE_DTYPE))
        assert after[100:600].tobytes() == empty_caches(500).tobytes()
        assert after[:100].tobytes() == filled[:100].tobytes()
        assert after[600:].tobytes() == filled[600:].tobytes()
    finally:
        for b in (ca, dp, dx, do):
            b.free()
