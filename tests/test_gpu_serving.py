"""GPU tests of the serving loop (b3amd_server_*): continuously batched world steps give
the oracle's manifolds and feature caches bit for bit with each world's caches chained
across its steps (as every b3ConvexContact keeps its m_cache), a world's results do not
depend on what it is batched with, world admission waits for slots in arrival order, and
destroying a world while another thread waits on its step is safe."""
import sys
import threading
import time

import numpy as np
import pytest

import bounce_amd
from bounce_amd.types import empty_caches
from bounce_amd.workloads import gen_serving_worlds, world_pair_inputs
from util import assert_caches_equal, assert_manifolds_equal, defined_hull_cache

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def serving(lib):
    pool, shapes, worlds = gen_serving_worlds(lib, 31, 24, bodies_min=10, bodies_max=300, steps=3)
    return pool, shapes, worlds


def body_sweeps(world, step, seed):
    """Per body (orientation0, orientation): small rotations about random axes, so the
    cached face rebuild's orientation gate (collide_hulls_cache.cpp:117-152) both passes
    and fails across the bodies."""
    rng = np.random.default_rng(seed * 1000 + step)
    n = len(world["body_shapes"])
    sw = np.zeros((n, 2, 4), np.float32)
    for k in range(2):
        ax = rng.normal(size=(n, 3))
        ax /= np.linalg.norm(ax, axis=1, keepdims=True)
        half = rng.uniform(0.0, 0.08, n) / 2.0
        sw[:, k, :3] = ax * np.sin(half)[:, None]
        sw[:, k, 3] = np.cos(half)
    return sw


def pair_sweeps(world, sw):
    bp = world["pairs"].astype(np.int64)
    return np.ascontiguousarray(np.concatenate([sw[bp[:, 0]], sw[bp[:, 1]]], 1))


def oracle_step(oracle, pool, shapes, world, step, caches, sw=None):
    pairs, xfs = world_pair_inputs(world, step)
    return oracle.collide(pool, shapes, pairs, xfs, sweeps=None if sw is None else pair_sweeps(world, sw),
                          caches=caches)


def test_gpu_serving_equals_oracle(lib, oracle, serving):
    """Three steps of 24 worlds, all submitted together each step, even worlds with body
    sweeps and odd worlds without: every world's manifolds and caches equal the oracle's
    chain for that world."""
    say("case: serving vs oracle, caches chained per world")
    pool, shapes, worlds = serving
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        caches = [empty_caches(len(w["pairs"])) for w in worlds]
        with bounce_amd.Server(ctx, 1 << 14, 1 << 14, 1 << 16) as srv:
            ids = [srv.create_world(w["body_shapes"], w["pairs"]) for w in worlds]
            for step in range(3):
                sws = [body_sweeps(w, step, k) if k % 2 == 0 else None for k, w in enumerate(worlds)]
                tickets = [srv.submit(i, w["xfs"][step], sw) for i, w, sw in zip(ids, worlds, sws)]
                for t in tickets:
                    assert srv.wait(t) >= 0.0
                for k, (i, w) in enumerate(zip(ids, worlds)):
                    want = oracle_step(oracle, pool, shapes, w, step, caches[k], sws[k])
                    what = "world %d step %d" % (k, step)
                    assert_manifolds_equal(srv.read_manifolds(i), want, what)
                    assert_caches_equal(defined_hull_cache(srv.read_caches(i)), defined_hull_cache(caches[k]), what)
            st = srv.stats()
            assert st["requests"] == 3 * len(worlds)
            assert st["batches"] < st["requests"]  # steps of several worlds shared batches
    # Every hull-hull contact's feature cache holds a feature after its first step (sphere
    # pairs do not use it).
    from bounce_amd.types import HULL
    for w, c in zip(worlds, caches):
        t = shapes["type"][w["body_shapes"][w["pairs"].astype(np.int64)]]
        hh = (t[:, 0] == HULL) & (t[:, 1] == HULL)
        assert (c["state"][hh] != 2).all() and hh.any()


def test_gpu_serving_alone_vs_mix(lib, serving):
    """World 0 stepped alone and stepped inside a mix of 23 other worlds (several steps of
    each world queued at once, so batches overlap on the device and wait for each other):
    identical manifolds and caches at every step."""
    say("case: serving alone vs mix")
    pool, shapes, worlds = serving
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        with bounce_amd.Server(ctx, 1 << 14, 1 << 14, 1 << 16) as srv:
            srv.configure(batch_pairs=2048, max_inflight=3)
            w0 = worlds[0]
            alone = srv.create_world(w0["body_shapes"], w0["pairs"])
            got_alone = []
            for step in range(3):
                srv.wait(srv.submit(alone, w0["xfs"][step], body_sweeps(w0, step, 0)))
                got_alone.append((srv.read_manifolds(alone), srv.read_caches(alone)))
            srv.destroy_world(alone)
            others = [srv.create_world(w["body_shapes"], w["pairs"]) for w in worlds[1:]]
            mixed = srv.create_world(w0["body_shapes"], w0["pairs"])
            tickets = []
            for step in range(3):  # all three steps queued before any completes
                tickets += [srv.submit(i, w["xfs"][step]) for i, w in zip(others[:11], worlds[1:12])]
                tickets.append(srv.submit(mixed, w0["xfs"][step], body_sweeps(w0, step, 0)))
                tickets += [srv.submit(i, w["xfs"][step]) for i, w in zip(others[11:], worlds[12:])]
            for t in tickets:
                srv.wait(t)
            # Only the last step's state is readable now; replay to compare each step.
            assert_manifolds_equal(srv.read_manifolds(mixed), got_alone[2][0], "step 2")
            assert_caches_equal(srv.read_caches(mixed), got_alone[2][1], "step 2")
            srv.destroy_world(mixed)
            mixed = srv.create_world(w0["body_shapes"], w0["pairs"])
            for step in range(3):
                ts = [srv.submit(i, w["xfs"][step]) for i, w in zip(others, worlds[1:])]
                ts.insert(7, srv.submit(mixed, w0["xfs"][step], body_sweeps(w0, step, 0)))
                for t in ts:
                    srv.wait(t)
                assert_manifolds_equal(srv.read_manifolds(mixed), got_alone[step][0], "step %d" % step)
                assert_caches_equal(srv.read_caches(mixed), got_alone[step][1], "step %d" % step)


def test_gpu_serving_admission(lib, serving):
    """A world that does not fit the slab is refused by b3amd_world_create and by
    b3amd_world_admit once its time is up; an admission waiting in another thread gets the
    slots when a world leaves, before a later admission; a slab beyond the context's HBM
    budget is refused."""
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
            t0 = time.perf_counter()
            with pytest.raises(bounce_amd.BounceError) as e:
                srv.admit_world(w["body_shapes"], w["pairs"], 50.0)
            assert e.value.code == bounce_amd.ERR_BUDGET and time.perf_counter() - t0 >= 0.045
            got = {}

            def admit(tag, delay):
                time.sleep(delay)
                got[tag] = (srv.admit_world(w["body_shapes"], w["pairs"], 20000.0), time.perf_counter())

            th = [threading.Thread(target=admit, args=("first", 0.0)), threading.Thread(target=admit, args=("second", 0.1))]
            for t in th:
                t.start()
            time.sleep(0.3)
            assert not got  # both wait: no room
            srv.wait(srv.submit(a, w["xfs"][0]))
            srv.destroy_world(a)
            th[0].join(10)
            assert "first" in got and "second" not in got  # arrival order: only one fits
            srv.wait(srv.submit(got["first"][0], w["xfs"][0]))
            srv.destroy_world(got["first"][0])
            th[1].join(10)
            assert "second" in got
            srv.wait(srv.submit(got["second"][0], w["xfs"][1]))
        with pytest.raises(bounce_amd.BounceError) as e:
            bounce_amd.Server(ctx, 1 << 14, 1 << 14, (1 << 32) - 1)  # slab beyond the 64 GB budget
        assert e.value.code == bounce_amd.ERR_BUDGET


def test_gpu_serving_destroy_while_waiting(lib, serving):
    """One thread waits on a world's last step while another destroys the world: the
    destroy returns after the step completes, the waiter gets its latency, and the world's
    slots are reusable."""
    say("case: serving destroy while another thread waits")
    pool, shapes, worlds = serving
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        with bounce_amd.Server(ctx, 1 << 14, 1 << 14, 1 << 16) as srv:
            for rep in range(20):
                ids = [srv.create_world(w["body_shapes"], w["pairs"]) for w in worlds[:6]]
                tickets = [srv.submit(i, w["xfs"][s % 3]) for s in range(3) for i, w in zip(ids, worlds[:6])]
                lat = {}
                waiter = threading.Thread(target=lambda: lat.update(v=[srv.wait(t) for t in tickets]))
                waiter.start()
                for i in ids:
                    srv.destroy_world(i)
                waiter.join(30)
                assert not waiter.is_alive() and len(lat["v"]) == len(tickets) and min(lat["v"]) > 0
                with pytest.raises(bounce_amd.BounceError):
                    srv.submit(ids[0], worlds[0]["xfs"][0])


def test_gpu_serving_full_size(lib, oracle):
    """The serving workload at full size: 4,096 worlds of 10-2,000 bodies admitted at once
    and stepped twice through the server (every world's step submitted before any is
    waited on); the manifolds and caches of every 64th world equal the oracle's chain, and
    every world's manifolds are well formed."""
    say("case: serving full size, 4,096 worlds")
    pool, shapes, worlds = gen_serving_worlds(lib, 4096, 4096, steps=2)
    contacts = sum(len(w["pairs"]) for w in worlds)
    sample = list(range(0, len(worlds), 64))
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        mp = max(len(w["pairs"]) for w in worlds)
        mb = max(len(w["body_shapes"]) for w in worlds)
        with bounce_amd.Server(ctx, max(1 << 16, mp), max(1 << 17, mb), contacts) as srv:
            srv.configure(batch_pairs=1 << 16, max_inflight=3)
            ids = [srv.admit_world(w["body_shapes"], w["pairs"], 1000.0) for w in worlds]
            caches = {k: empty_caches(len(worlds[k]["pairs"])) for k in sample}
            for step in range(2):
                tickets = [srv.submit(i, w["xfs"][step]) for i, w in zip(ids, worlds)]
                for t in tickets:
                    srv.wait(t)
                for k in sample:
                    want = oracle_step(oracle, pool, shapes, worlds[k], step, caches[k])
                    assert_manifolds_equal(srv.read_manifolds(ids[k]), want, "world %d step %d" % (k, step))
                    assert_caches_equal(defined_hull_cache(srv.read_caches(ids[k])), defined_hull_cache(caches[k]),
                                        "world %d step %d" % (k, step))
            touching = 0
            for i in ids:
                m = srv.read_manifolds(i)
                assert m["pointCount"].max() <= 4 and (m["flags"] == 0).all()
                n = m["points"]["localNormal1"][m["pointCount"] > 0, 0]
                assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)
                touching += int((m["pointCount"] > 0).sum())
            assert 0.2 < touching / contacts < 0.5
            assert srv.stats()["requests"] == 2 * len(worlds)


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
            srv.configure(max_inflight=3)
            ids = [srv.create_world(w["body_shapes"], w["pairs"]) for w in worlds]
            lat = srv.replay(np.array(ids, np.uint32)[wsel], times, xfs)
            assert (lat > 0).all() and np.isfinite(lat).all()
            say("replay p50 %.3f ms p99 %.3f ms, %s" % (np.percentile(lat, 50), np.percentile(lat, 99), srv.stats()))
            assert srv.stats()["requests"] == n
