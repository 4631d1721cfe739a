"""GPU tests of the serving loop (b3amd_server_*): continuously batched world steps give
the batch API's manifolds bit for bit, a world's results do not depend on what it is
batched with, and world admission respects the slab capacity."""
import sys

import numpy as np
import pytest

import bounce_amd
from bounce_amd.types import empty_caches
from bounce_amd.workloads import gen_serving_worlds, world_pair_inputs
from util import assert_manifolds_equal

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def serving(lib):
    pool, shapes, worlds = gen_serving_worlds(lib, 31, 24, bodies_min=10, bodies_max=300, steps=3)
    return pool, shapes, worlds


def test_gpu_serving_equals_batch_api(lib, serving):
    """Three steps of 24 worlds, all submitted together each step: every world's manifolds
    equal b3amd_collide on the same contacts with persistent caches."""
    say("case: serving vs batch API")
    pool, shapes, worlds = serving
    with bounce_amd.Context(0) as ctx, bounce_amd.Context(0) as ref:
        for c in (ctx, ref):
            c.upload_hulls(pool)
            c.upload_shapes(shapes)
        caches = [empty_caches(len(w["pairs"])) for w in worlds]
        with bounce_amd.Server(ctx, 1 << 14, 1 << 14, 1 << 16) as srv:
            ids = [srv.create_world(w["body_shapes"], w["pairs"]) for w in worlds]
            for step in range(3):
                tickets = [srv.submit(i, w["xfs"][step]) for i, w in zip(ids, worlds)]
                for t in tickets:
                    assert srv.wait(t) >= 0.0
                for k, (i, w) in enumerate(zip(ids, worlds)):
                    pairs, xfs = world_pair_inputs(w, step)
                    want = ref.collide(pairs, xfs, caches=caches[k])
                    assert_manifolds_equal(srv.read_manifolds(i), want, "world %d step %d" % (k, step))
            st = srv.stats()
            assert st["requests"] == 3 * len(worlds)
            assert st["batches"] < st["requests"]  # steps of several worlds shared batches


def test_gpu_serving_alone_vs_mix(lib, serving):
    """World 0 stepped alone and stepped inside a mix of 23 other worlds: identical
    manifolds at every step."""
    say("case: serving alone vs mix")
    pool, shapes, worlds = serving
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        with bounce_amd.Server(ctx, 1 << 14, 1 << 14, 1 << 16) as srv:
            w0 = worlds[0]
            alone = srv.create_world(w0["body_shapes"], w0["pairs"])
            got_alone = []
            for step in range(3):
                srv.wait(srv.submit(alone, w0["xfs"][step]))
                got_alone.append(srv.read_manifolds(alone))
            srv.destroy_world(alone)
            others = [srv.create_world(w["body_shapes"], w["pairs"]) for w in worlds[1:]]
            mixed = srv.create_world(w0["body_shapes"], w0["pairs"])
            for step in range(3):
                tickets = [srv.submit(i, w["xfs"][step]) for i, w in zip(others[:11], worlds[1:12])]
                tickets.append(srv.submit(mixed, w0["xfs"][step]))
                tickets += [srv.submit(i, w["xfs"][step]) for i, w in zip(others[11:], worlds[12:])]
                for t in tickets:
                    srv.wait(t)
                assert_manifolds_equal(srv.read_manifolds(mixed), got_alone[step], "step %d" % step)


def test_gpu_serving_admission(lib, serving):
    """A world that does not fit the slab is refused with B3AMD_ERR_BUDGET and admitted once
    another world has left."""
    say("case: serving admission")
    pool, shapes, worlds = serving
    w = max(worlds, key=lambda x: len(x["pairs"]))
    n = len(w["pairs"])
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        with bounce_amd.Server(ctx, 1 << 14, 1 << 14, n + n // 2) as srv:
            a = srv.create_world(w["body_shapes"], w["pairs"])
            with pytest.raises(bounce_amd.BounceError) as e:
                srv.create_world(w["body_shapes"], w["pairs"])
            assert e.value.code == bounce_amd.ERR_BUDGET
            srv.destroy_world(a)
            b = srv.create_world(w["body_shapes"], w["pairs"])
            srv.wait(srv.submit(b, w["xfs"][0]))
        with pytest.raises(bounce_amd.BounceError) as e:
            bounce_amd.Server(ctx, 1 << 14, 1 << 14, (1 << 32) - 1)  # slab beyond the 64 GB budget
        assert e.value.code == bounce_amd.ERR_BUDGET


def test_gpu_serving_replay(lib, serving):
    """The load generator: 2,000 Poisson arrivals over the 24 worlds; every request
    completes with a latency measured from its scheduled arrival."""
    say("case: serving replay")
    pool, shapes, worlds = serving
    rng = np.random.default_rng(3)
    n = 2000
    times = np.cumsum(rng.exponential(1.0 / 40000.0, n))
    wsel = rng.integers(0, len(worlds), n)
    steps = rng.integers(0, 3, n)
    xfs = [worlds[w]["xfs"][s] for w, s in zip(wsel, steps)]
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        with bounce_amd.Server(ctx, 1 << 14, 1 << 14, 1 << 16) as srv:
            ids = [srv.create_world(w["body_shapes"], w["pairs"]) for w in worlds]
            lat = srv.replay(np.array(ids, np.uint32)[wsel], times, xfs)
            assert (lat > 0).all() and np.isfinite(lat).all()
            say("replay p50 %.3f ms p99 %.3f ms, %s" % (np.percentile(lat, 50), np.percentile(lat, 99), srv.stats()))
            assert srv.stats()["requests"] == n
