"""GPU statistics counters (b3amd_get_counters) equal the oracle's counts of the
reference's globals b3_convexCalls, b3_convexCacheHits, b3_gjkCalls, b3_gjkIters and
b3_gjkMaxIters, which the CPU suite pins to the compiled reference."""
import sys

import numpy as np
import pytest

import bounce_amd
from bounce_amd.types import empty_caches
from util import hull_workload, mixed_workload, mesh_workload

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_gpu_counters_equal_oracle(lib, oracle):
    say("case: counters (warm hulls, mixed convex, mesh)")
    import oracle_lib
    with bounce_amd.Context(0) as ctx:
        ctx.set_counting(True)
        pool, shapes, pairs, xfs = hull_workload(lib, 77, 200, 3000)
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        cg, co = empty_caches(len(pairs)), empty_caches(len(pairs))
        rng = np.random.default_rng(9)
        x = xfs.copy()
        ctx.reset_counters()
        oracle.reset_counters()
        for step in range(3):
            ctx.collide(pairs, x, caches=cg)
            oracle.collide(pool, shapes, pairs, x, caches=co)
            assert ctx.counters() == oracle.counters(), "hull step %d" % step
            x[:, 1, 9:12] += rng.normal(0, 0.003, size=(len(pairs), 3)).astype(np.float32)
        assert ctx.counters()["convex_cache_hits"] > 0

        mpool, mshapes, mpairs, mxfs = mixed_workload(lib, 23, 3000)
        ctx.upload_hulls(mpool)
        ctx.upload_shapes(mshapes)
        ctx.reset_counters()
        oracle.reset_counters()
        ctx.collide(mpairs, mxfs)
        oracle.collide(mpool, mshapes, mpairs, mxfs)
        assert ctx.counters() == oracle.counters()
        assert ctx.counters()["gjk_calls"] > 0

        pool, shapes, v, t, mx = mesh_workload(lib, 11, 32, 32, 300)
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        mid = ctx.upload_mesh(v, t)
        nodes = ctx.mesh_tree(mid)
        sa = np.arange(len(shapes), dtype=np.uint32)
        ctx.reset_counters()
        oracle.reset_counters()
        counts = ctx.collide_mesh(mid, sa, mx)[0]
        assert (counts <= ctx.MESH_TRIANGLE_CAP).all()  # every triangle collided on both sides
        oracle_lib.collide_mesh(oracle.lib, "b3o_", pool, shapes, v, t, nodes, sa, mx)
        assert ctx.counters() == oracle.counters()

        # Counting off: the totals stay where they are.
        ctx.set_counting(False)
        ctx.reset_counters()
        ctx.collide_mesh(mid, sa, mx)
        assert all(v == 0 for v in ctx.counters().values())
