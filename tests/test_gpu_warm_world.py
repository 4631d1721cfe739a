"""GPU tests of b3amd_update_contacts on the warm-world workload: world steps from
device-resident body state, with each contact's feature cache kept on the device from
step to step, give the oracle's manifolds and caches bit for bit at every step; the
full-size world (200,000 hulls) agrees with the compiled reference on a strided sample."""
import sys

import numpy as np
import pytest

import bounce_amd
from bounce_amd.types import FLAG_INVALID_PAIR, MANIFOLD_DTYPE, empty_caches
from bounce_amd.workloads import gen_warm_world, warm_world_pair_inputs, warm_world_state
from util import (HULL_CACHE_FIELDS, WarmWorldDevice, assert_caches_equal, assert_manifolds_equal,
                  defined_hull_cache, warm_world_chain)

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_world(lib):
    return gen_warm_world(lib, 17, n_hulls=900, columns=6, half_width=5.0)


def test_gpu_update_contacts_warm_chain(lib, oracle, small_world):
    """Twelve steps (the first cold, the rest warm with sweeps): manifolds and the defined
    cache fields equal the oracle's chain at every step; the warm steps hit the caches."""
    w = small_world
    say("case: warm world chain, %d bodies, %d contacts, 12 steps" % (len(w["body_shapes"]), len(w["contacts"])))
    steps = 12
    want = warm_world_chain(oracle, w, steps, keep=range(steps))
    with bounce_amd.Context(0) as ctx:
        dev = WarmWorldDevice(ctx, w)
        ctx.set_counting(True)
        for s in range(steps):
            ctx.reset_counters()
            dev.step(s)
            out, caches = dev.read()
            assert_manifolds_equal(out, want[s][0], "step %d" % s)
            assert_caches_equal(defined_hull_cache(caches), defined_hull_cache(want[s][1]), "step %d caches" % s,
                                HULL_CACHE_FIELDS)
            cnt = ctx.counters()
            assert cnt["convex_calls"] == len(w["contacts"])
            if s == 0:
                assert cnt["convex_cache_hits"] == 0
            else:
                assert cnt["convex_cache_hits"] > 0.5 * len(w["contacts"])
        assert (out["pointCount"] > 0).any()


def test_gpu_update_contacts_equals_batch_api(lib, small_world):
    """Three steps without sweeps: b3amd_update_contacts gives what b3amd_collide gives on
    the per-contact inputs gathered on the host, caches included."""
    w = small_world
    say("case: update_contacts vs collide, no sweeps")
    with bounce_amd.Context(0) as ctx, bounce_amd.Context(0) as ref:
        ref.upload_hulls(w["pool"])
        ref.upload_shapes(w["shapes"])
        dev = WarmWorldDevice(ctx, w)
        caches = empty_caches(len(w["contacts"]))
        for s in range(3):
            dev.step(s, sweeps=False)
            out, got_caches = dev.read()
            xfs, sweeps = warm_world_state(w, s)
            pairs, pxf, _ = warm_world_pair_inputs(w, xfs, sweeps)
            want = ref.collide(pairs, pxf, caches=caches)
            assert_manifolds_equal(out, want, "step %d" % s)
            assert_caches_equal(defined_hull_cache(got_caches), defined_hull_cache(caches), "step %d caches" % s,
                                HULL_CACHE_FIELDS)


def test_gpu_update_contacts_invalid_body(lib, small_world):
    """A contact naming a body outside the world is flagged invalid with no points; the
    other contacts are unaffected."""
    w = small_world
    say("case: update_contacts invalid body index")
    B = len(w["body_shapes"])
    contacts = w["contacts"][:64].copy()
    contacts[5] = (3, B)
    contacts[9] = (B + 7, 2)
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(w["pool"])
        ctx.upload_shapes(w["shapes"])
        d_shapes, d_contacts = ctx.alloc(4 * B), ctx.alloc(8 * len(contacts))
        d_xfs, d_out = ctx.alloc(48 * B), ctx.alloc(MANIFOLD_DTYPE.itemsize * len(contacts))
        d_shapes.upload(w["body_shapes"])
        d_contacts.upload(contacts)
        xfs, sweeps = warm_world_state(w, 0)
        d_xfs.upload(xfs)
        ctx.update_contacts(B, d_shapes, d_xfs, len(contacts), d_contacts, d_out)
        ctx.synchronize()
        out = d_out.download(np.zeros(len(contacts), MANIFOLD_DTYPE))
        ok = np.ones(len(contacts), bool)
        ok[[5, 9]] = False
        pairs, pxf, _ = warm_world_pair_inputs(w, xfs, sweeps, contacts[ok])
        want = ctx.collide(pairs, pxf)
    assert (out["flags"][~ok] & FLAG_INVALID_PAIR).all()
    assert (out["pointCount"][~ok] == 0).all()
    assert_manifolds_equal(out[ok], want, "valid contacts")


def test_gpu_warm_world_full_size_sampled(lib, reference):
    """The 200,000-hull world (~2.8 M contacts) for four steps on the device; every 499th
    contact's manifold and cache after step 3 equal the compiled reference's chain of the
    same contacts."""
    say("case: full-size warm world, 4 steps")
    w = gen_warm_world(lib, 11)
    steps = 4
    idx = np.arange(0, len(w["contacts"]), 499)
    want_out, want_caches = warm_world_chain(reference, w, steps, idx)[steps - 1]
    with bounce_amd.Context(0) as ctx:
        dev = WarmWorldDevice(ctx, w)
        for s in range(steps):
            dev.step(s)
        out, caches = dev.read()
    assert_manifolds_equal(out[idx], want_out, "sampled contacts")
    assert_caches_equal(defined_hull_cache(caches[idx]), defined_hull_cache(want_caches), "sampled caches",
                        HULL_CACHE_FIELDS)
    assert (out["flags"] == 0).all()
    assert 0.03 < (out["pointCount"] > 0).mean() < 0.3
