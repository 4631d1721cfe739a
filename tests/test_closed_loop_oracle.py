"""The closed-loop drop-in on the CPU: the reference's own b3World steps the box pyramid
(20 levels, 210 unit boxes on a 100 x 1 x 100 m static box, 1/60 s, 8 velocity and 2
position iterations) twice, once with its own narrowphase and once with every
b3ConvexContact::Collide (convex_contact.cpp:52-63) served from a batch computed before
the step from the contacts UpdateContacts will collide (contact_manager.cpp:167-231), the
feature caches living in the contacts as in the reference.  Here the batch comes from the
oracle; the GPU test of the same loop takes it from the HIP library.  Every body's
transform, sweep and velocities must be bit-identical after every step.

The loop world is the compiled reference with b3CollideShapeAndShape wrapped at link time
(oracle/Makefile), so the Update tail (contact.cpp:63-140), islands and solver run
unchanged; a served call checks that the transforms it is handed are the ones the batch
was computed from."""
import numpy as np
import pytest

from util import pyramid_shapes

STEPS = 300


def test_closed_loop_with_oracle_manifolds(reference, oracle):
    from oracle_lib import RefPyramid
    plain, loop = RefPyramid(reference.lib), RefPyramid(reference.lib)
    try:
        pool = plain.hulls()
        shapes = pyramid_shapes(plain.body_count)
        assert np.array_equal(plain.state(), loop.state())
        served = 0
        for s in range(STEPS):
            g = loop.gather()
            caches = g["caches"]
            out = oracle.collide(pool, shapes, g["pairs"], g["xfs"], sweeps=g["sweeps"], caches=caches)
            stats = loop.loop_step(g, out, caches)
            plain.step()
            assert stats == (len(g["ids"]), 0, 0), "step %d: served, missed, mismatched = %s" % (s, stats)
            served += stats[0]
            a, b = plain.state(), loop.state()
            assert a.tobytes() == b.tobytes(), "step %d: body %s differs" % (
                s, np.nonzero((a.view(np.uint32) != b.view(np.uint32)).any(axis=1))[0][:5])
        assert served > 20000
        # The pyramid has settled into resting contact, not fallen apart.
        st = loop.state()
        assert np.abs(st[1:, 26:32]).max() < 0.5
    finally:
        plain.close()
        loop.close()


@pytest.mark.parametrize("mistake", ["point", "key"])
def test_closed_loop_sees_a_planted_mistake(reference, oracle, mistake):
    """The check is not vacuous: in one step of the served batch, one manifold point moved
    by 10 micrometres, or one feature key with a bit flipped (the point then does not
    persist for warm starting, manifold.cpp:34-55), makes the loop world leave the plain
    world."""
    from oracle_lib import RefPyramid
    plain, loop = RefPyramid(reference.lib), RefPyramid(reference.lib)
    try:
        pool = plain.hulls()
        shapes = pyramid_shapes(plain.body_count)
        differ = False
        for s in range(240):
            g = loop.gather()
            caches = g["caches"]
            out = oracle.collide(pool, shapes, g["pairs"], g["xfs"], sweeps=g["sweeps"], caches=caches)
            if s == 200:
                k = int(np.nonzero(out["pointCount"] > 0)[0][0])
                if mistake == "point":
                    out["points"]["localPoint1"][k, 0, 0] += np.float32(1e-5)
                else:
                    out["points"]["key"][k, 0, 0] ^= np.uint32(1)
            loop.loop_step(g, out, caches)
            plain.step()
            if s < 200:
                assert plain.state().tobytes() == loop.state().tobytes()
            else:
                differ = differ or plain.state().tobytes() != loop.state().tobytes()
        assert differ
    finally:
        plain.close()
        loop.close()


def test_plain_world_unaffected_by_wrap(reference):
    """Outside a loop step the wrapped call reaches the reference's own b3CollideShapeAndShape:
    stepping with the recording driver equals plain stepping in another world."""
    from oracle_lib import RefPyramid
    a, b = RefPyramid(reference.lib), RefPyramid(reference.lib)
    try:
        for s in range(10):
            r = a.step()
            b.step()
            assert len(r["ids"]) > 0 or s == 0
            assert a.state().tobytes() == b.state().tobytes()
    finally:
        a.close()
        b.close()
