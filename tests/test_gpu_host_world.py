"""The whole b3World::Step with no reference code: the library's host step (islands and
the sequential-impulse solver, restated from world.cpp, island.cpp and contact_solver.cpp)
around the device broadphase, contact manager and narrowphase, stepping the box pyramid
(20 levels, 210 unit boxes of density 1 on a static 100 x 1 x 100 m box, 1/60 s, 8 velocity
and 2 position iterations, 300 steps).  Every body's transform and sweep orientations after
every step equal the reference world's (the trace fixture written by the reference's own
b3World), bit for bit; where the reference is built, its live world also pins the sweep
centres and the velocities."""
import numpy as np
import pytest

import bounce_amd
import oracle_lib
from util import load_golden, PYRAMID_STEPS

pytestmark = pytest.mark.gpu

LEVELS = 20


def build_pyramid(ctx, t):
    from bounce_amd.types import SHAPE_DTYPE, HULL
    ctx.upload_hulls(t["pool"])       # hull 0: the 100 x 1 x 100 ground box, hull 1: the unit box
    shapes = np.zeros(2, SHAPE_DTYPE)
    shapes["type"] = HULL
    shapes["hull"] = [0, 1]
    ctx.upload_shapes(shapes)
    w = ctx.host_world()
    g = w.create_body(type=0)
    w.create_shape(g, 0)
    f32 = np.float32
    for l in range(LEVELS):
        for j in range(LEVELS - l):
            # the reference's placement in float32 (oracle/ref_driver.cpp: b3ref_pyramid_create)
            x = f32(2.1) * f32(j) + f32(l) * f32(1.0) - f32(0.5) * f32(LEVELS) * f32(2.0)
            b = w.create_body(type=2, position=(x, f32(3.0) * f32(l + 1), 0.0))
            w.create_shape(b, 1, density=1.0)
    return w


def test_gpu_host_world_pyramid_equals_reference():
    t = load_golden("pyramid_trace.npz")
    ref = oracle_lib.reference()
    live = oracle_lib.RefPyramid(ref.lib) if ref is not None else None
    print("host world: pyramid of %d levels, %d steps" % (LEVELS, PYRAMID_STEPS), flush=True)
    with bounce_amd.Context(0) as ctx:
        w = build_pyramid(ctx, t)
        try:
            st = w.bodies()
            assert np.array_equal(st["xf"].view(np.uint32), t["body_xfs"][0].view(np.uint32))
            if live is not None:
                rs = live.state()
                assert np.array_equal(st["linear_velocity"].view(np.uint32), rs[:, 26:29].view(np.uint32))
            for s in range(PYRAMID_STEPS - 1):
                w.step(1.0 / 60.0, 8, 2)
                st = w.bodies()
                got_xf = st["xf"]
                got_q = np.stack([st["sweep"]["orientation0"], st["sweep"]["orientation"]], 1)
                bad = np.nonzero((got_xf.view(np.uint32) != t["body_xfs"][s + 1].view(np.uint32)).any(1))[0]
                assert len(bad) == 0, "step %d: %d bodies differ, first %d" % (s, len(bad), bad[0])
                assert np.array_equal(got_q.view(np.uint32), t["body_sweeps"][s + 1].view(np.uint32)), "step %d" % s
                if live is not None:
                    live.step()
                    rs = live.state()
                    assert np.array_equal(st["sweep"]["worldCenter"].view(np.uint32), rs[:, 12:15].view(np.uint32))
                    assert np.array_equal(st["linear_velocity"].view(np.uint32), rs[:, 26:29].view(np.uint32))
                    assert np.array_equal(st["angular_velocity"].view(np.uint32), rs[:, 29:32].view(np.uint32))
            assert w.contact_count() > 300
            # the pyramid has settled: every box rests (speeds below 1 cm/s)
            assert np.abs(st["linear_velocity"][1:]).max() < 0.05
        finally:
            w.close()
            if live is not None:
                live.close()
