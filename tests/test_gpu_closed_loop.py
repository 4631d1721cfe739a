"""The closed-loop drop-in on the GPU: the reference's own b3World steps the box pyramid
(20 levels, 210 unit boxes, 300 steps at 60 Hz) with every b3ConvexContact::Collide
(convex_contact.cpp:52-63) served from one HIP batch per step (b3amd_collide), computed
for the contacts UpdateContacts will collide (contact_manager.cpp:167-231) from the loop
world's own transforms, sweeps and the feature caches its contacts hold.  The Update tail
(contact.cpp:63-140), islands and solver are the reference's unchanged code.  Every body's
transform, sweep and velocities equal, bit for bit and after every step, those of a second
reference world stepping with its own narrowphase."""
import sys

import numpy as np
import pytest

from util import pyramid_shapes

pytestmark = pytest.mark.gpu

STEPS = 300


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_gpu_closed_loop_pyramid(ctx, reference):
    from oracle_lib import RefPyramid
    say("case: closed-loop pyramid, 300 steps with HIP manifolds")
    plain, loop = RefPyramid(reference.lib), RefPyramid(reference.lib)
    try:
        ctx.upload_hulls(plain.hulls())
        ctx.upload_shapes(pyramid_shapes(plain.body_count))
        served = points = 0
        for s in range(STEPS):
            g = loop.gather()
            caches = g["caches"]
            out = ctx.collide(g["pairs"], g["xfs"], sweeps=g["sweeps"], caches=caches)
            stats = loop.loop_step(g, out, caches)
            plain.step()
            assert stats == (len(g["ids"]), 0, 0), "step %d: served, missed, mismatched = %s" % (s, stats)
            served += stats[0]
            points += int(out["pointCount"].sum())
            a, b = plain.state(), loop.state()
            assert a.tobytes() == b.tobytes(), "step %d: bodies %s differ" % (
                s, np.nonzero((a.view(np.uint32) != b.view(np.uint32)).any(axis=1))[0][:5])
        assert served > 50000 and points > 300000
        assert np.abs(loop.state()[1:, 26:32]).max() < 0.5  # settled, not scattered
    finally:
        plain.close()
        loop.close()
