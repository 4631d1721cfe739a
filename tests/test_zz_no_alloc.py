"""b3amd_collide_device_noalloc, called through the C ABI, gives b3amd_collide_device's
manifolds and caches bit for bit on two calls in a row (the second smaller, so stale
contents of the shared scratch from the first must not leak into it), allocates nothing
(the context's device memory stays the same) and refuses a batch above its maximum."""
import ctypes
import sys

import numpy as np
import pytest

import bounce_amd
from bounce_amd.types import MANIFOLD_DTYPE, HullPool, empty_caches
from util import assert_caches_equal, assert_manifolds_equal, hull_workload, mixed_workload

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc(lib):
    f = lib.b3amd_collide_device_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32] + [ctypes.c_void_p] * 5
    return f


def _pool_and_pairs(lib):
    """Hull pairs of 8-64 points and mixed sphere/capsule/hull pairs in one shape table."""
    hp, hs, hpairs, hx = hull_workload(lib, 4242, 200, 24000)
    mp, ms, mpairs, mx = mixed_workload(lib, 4343, 6000)
    pool = HullPool.concat([hp.hull(h) for h in range(hp.count)] + [mp.hull(h) for h in range(mp.count)])
    ms = ms.copy()
    ms["hull"] += hp.count
    shapes = np.concatenate([hs, ms])
    mpairs = mpairs + np.uint32(len(hs))
    pairs = np.concatenate([hpairs, mpairs])
    xfs = np.concatenate([hx, mx])
    perm = np.random.default_rng(1).permutation(len(pairs))
    return pool, shapes, pairs[perm], xfs[perm]


def test_gpu_noalloc_equals_collide_device(lib):
    say("case: no-allocation entry point vs b3amd_collide_device")
    f = _noalloc(lib)
    pool, shapes, pairs, xfs = _pool_and_pairs(lib)
    rng = np.random.default_rng(7)
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        N = len(pairs)
        bufs = {k: ctx.alloc(b) for k, b in (("p", pairs.nbytes), ("x", xfs.nbytes), ("s", N * 64),
                                              ("c1", N * 36), ("c2", N * 36),
                                              ("o1", N * MANIFOLD_DTYPE.itemsize), ("o2", N * MANIFOLD_DTYPE.itemsize))}
        try:
            for call, n in enumerate((N, N // 7)):  # the second call is smaller
                x = xfs[:n].copy()
                x[:, 1, 9:12] += rng.normal(0, 0.002, (n, 3)).astype(np.float32)
                sw = np.zeros((n, 4, 4), np.float32)
                sw[:, :, 3] = 1.0
                tilt = rng.uniform(0, 0.05, n).astype(np.float32)
                sw[:, 3, 0], sw[:, 3, 3] = tilt, np.sqrt(1 - tilt * tilt)
                bufs["p"].upload(np.ascontiguousarray(pairs[:n]))
                bufs["x"].upload(x)
                bufs["s"].upload(sw)
                for with_caches in (False, True):
                    caches = empty_caches(n)
                    if call == 1:  # warm caches from a first pass, so the cached paths run
                        ctx.collide(pairs[:n], x, caches=caches)
                    bufs["c1"].upload(caches)
                    bufs["c2"].upload(caches)
                    dc1 = bufs["c1"] if with_caches else None
                    dc2 = bufs["c2"].ptr if with_caches else None
                    dsw = bufs["s"].ptr if with_caches else None
                    ctx.collide_device(n, bufs["p"], bufs["x"], bufs["o1"], d_sweeps=bufs["s"] if with_caches else None,
                                       d_caches=dc1)
                    ctx.synchronize()
                    held = ctx.memory_in_use
                    rc = f(ctx.handle, n, bufs["p"].ptr, bufs["x"].ptr, dsw, dc2, bufs["o2"].ptr)
                    assert rc == 0, lib.b3amd_last_error()
                    ctx.synchronize()
                    assert ctx.memory_in_use == held
                    want = bufs["o1"].download(np.zeros(n, MANIFOLD_DTYPE))
                    got = bufs["o2"].download(np.zeros(n, MANIFOLD_DTYPE))
                    what = "call %d (n=%d, caches=%s)" % (call, n, with_caches)
                    assert_manifolds_equal(got, want, what)
                    if with_caches:
                        assert_caches_equal(bufs["c2"].download(empty_caches(n)), bufs["c1"].download(empty_caches(n)),
                                            what)
                    assert (want["pointCount"] > 0).sum() > n // 4
            rc = f(ctx.handle, 262145, bufs["p"].ptr, bufs["x"].ptr, None, None, bufs["o2"].ptr)
            assert rc == bounce_amd.ERR_ARGUMENT
        finally:
            for b in bufs.values():
                b.free()
