"""The box pyramid (20 levels, 210 unit boxes on a static 100 x 1 x 100 m box) as the
reference's own b3World steps it, 300 steps at 60 Hz: per step, the manifolds the HIP
narrowphase computes for the contacts b3ContactManager::UpdateContacts collides
(contact_manager.cpp:167-231) are compared pair by pair with the reference's, bit-exact,
with the feature caches carried from step to step by the device path itself."""
import sys

import numpy as np
import pytest

from util import (HULL_CACHE_FIELDS, PYRAMID_FULL_STEPS, PYRAMID_STEPS, CacheChain, assert_caches_equal,
                  assert_manifolds_equal, defined_hull_cache, load_golden, pyramid_full_outputs, pyramid_inputs, pyramid_shapes,
                  step_digest)

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_gpu_pyramid_golden(ctx):
    say("case: pyramid trace fixture, 300 steps")
    t = load_golden("pyramid_trace.npz")
    ctx.upload_hulls(t["pool"])
    ctx.upload_shapes(pyramid_shapes(t["body_xfs"].shape[1]))
    chain = CacheChain()
    points = 0
    for s in range(PYRAMID_STEPS):
        ids, pairs, xfs, sweeps = pyramid_inputs(t, s)
        caches = chain.gather(ids)
        out = ctx.collide(pairs, xfs, sweeps=sweeps, caches=caches)
        chain.scatter(ids, caches)
        points += int(out["pointCount"].sum())
        if s in PYRAMID_FULL_STEPS:
            want_out, want_caches = pyramid_full_outputs(t, s)
            assert_manifolds_equal(out, want_out, "step %d" % s)
            assert_caches_equal(caches, want_caches, "step %d" % s, HULL_CACHE_FIELDS)
        assert np.array_equal(step_digest(out, caches), t["digests"][s]), "step %d differs" % s
    assert points > 300000  # the pyramid settles into ~1,600 points per step


def test_gpu_pyramid_equals_reference_world(ctx, reference):
    """Live against the compiled reference world: each step's inputs and the reference's
    caches go to the device; manifolds and caches must equal the reference's; the device's
    own chained caches must equal the reference's caches before every step."""
    from oracle_lib import RefPyramid
    say("case: pyramid against the reference world, 300 steps")
    world = RefPyramid(reference.lib)
    ctx.upload_hulls(world.hulls())
    ctx.upload_shapes(pyramid_shapes(world.body_count))
    chain = CacheChain()
    for s in range(PYRAMID_STEPS):
        r = world.step()
        ids = r["ids"].astype(np.int64)
        assert_caches_equal(defined_hull_cache(chain.gather(ids)), defined_hull_cache(r["caches_in"]),
                            "chained caches before step %d" % s)
        caches = r["caches_in"].copy()
        out = ctx.collide(r["pairs"], r["xfs"], sweeps=r["sweeps"], caches=caches)
        assert_manifolds_equal(out, r["out"], "step %d" % s)
        assert_caches_equal(caches, r["caches_out"], "step %d" % s, HULL_CACHE_FIELDS)
        chain.scatter(ids, caches)
    world.close()
