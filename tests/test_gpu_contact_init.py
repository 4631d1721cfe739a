"""GPU parity of the step after the narrowphase (b3amd_contact_init) against the oracle:
warm-start matching (b3Manifold::Initialize(old), manifold.cpp:34-55, in contact.cpp:63-91's
loop) and b3ContactSolver::InitializeConstraints (contact_solver.cpp:70-276), bit for bit;
and a device pipeline where the narrowphase's manifolds feed it straight from HBM."""
import ctypes
import sys

import numpy as np
import pytest

from bounce_amd.types import (IMPULSES_DTYPE, MANIFOLD_DTYPE, VELOCITY_MANIFOLD_DTYPE, empty_caches)
from util import contact_init_workload

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _args(w):
    return (w["bodies"], w["state"], w["material"], w["counts"], w["manifolds"], w["old_counts"], w["old_manifolds"],
            w["old_impulses"])


@pytest.mark.parametrize("M,warm", [(1, True), (3, True), (1, False), (3, False)])
def test_gpu_contact_init_equals_oracle(ctx, lib, oracle, M, warm):
    say("case: contact init M=%d warm=%s" % (M, warm))
    w = contact_init_workload(lib, oracle, 141 + M, 6000, M)
    gi, gv = ctx.contact_init(M, *_args(w), warm=warm)
    oi, ov = oracle.contact_init(M, *_args(w), warm=warm)
    assert gi.tobytes() == oi.tobytes(), "impulses"
    for j in range(M):
        g, o = np.ascontiguousarray(gv[:, j]), np.ascontiguousarray(ov[:, j])
        bad = np.nonzero(g.view(np.uint8).reshape(len(g), -1) != o.view(np.uint8).reshape(len(o), -1))
        assert len(bad[0]) == 0, "slot %d: %d contacts differ, first %d" % (j, len(set(bad[0])), bad[0][0])
    assert (ov["pointCount"] > 0).sum() > 1000


def test_gpu_narrowphase_to_constraints_on_device(ctx, lib, oracle):
    """Two steps: the narrowphase writes manifolds into one of two device buffers, and
    b3amd_contact_init matches them against the other (the previous step's) and builds the
    constraints, all on device-resident arrays; the result equals the oracle's chain."""
    say("case: narrowphase -> contact init on device")
    from util import mixed_workload
    rng = np.random.default_rng(5)
    n = 4000
    pool, shapes, pairs, xfs = mixed_workload(lib, 951, n, gap_mean=-0.02, gap_sigma=0.03)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    w = contact_init_workload(lib, oracle, 952, n, 1)  # bodies, state and materials
    steps = [xfs]
    x = xfs.copy()
    x[:, 1, 9:12] += rng.normal(0.0, 0.002, (n, 3)).astype(np.float32)
    steps.append(x)
    ms = MANIFOLD_DTYPE.itemsize
    d = {k: ctx.alloc(b) for k, b in (("pairs", pairs.nbytes), ("xfs", xfs.nbytes), ("caches", n * 36),
                                      ("m0", n * ms), ("m1", n * ms), ("counts", 4 * n), ("bodies", 8 * n),
                                      ("state", w["state"].nbytes), ("mat", w["material"].nbytes),
                                      ("imp0", n * 48), ("imp1", n * 48), ("vc", n * VELOCITY_MANIFOLD_DTYPE.itemsize))}
    try:
        d["pairs"].upload(pairs)
        d["caches"].upload(empty_caches(n))
        d["counts"].upload(np.ones(n, np.uint32))
        d["bodies"].upload(w["bodies"])
        d["state"].upload(w["state"])
        d["mat"].upload(w["material"])
        d["imp0"].upload(np.zeros(n, IMPULSES_DTYPE))
        f = lib.b3amd_contact_init
        f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32] + [ctypes.c_void_p] * 8 + [ctypes.c_int] + \
            [ctypes.c_void_p] * 2
        co = empty_caches(n)
        o_imp = np.zeros((n, 1), IMPULSES_DTYPE)
        o_man = None
        for step in range(2):
            cur, prev = ("m1", "m0") if step % 2 else ("m0", "m1")
            icur, iprev = ("imp1", "imp0") if step % 2 else ("imp0", "imp1")
            d["xfs"].upload(steps[step])
            ctx.collide_device(n, d["pairs"], d["xfs"], d[cur], d_caches=d["caches"])
            assert f(ctx.handle, n, 1, d["bodies"].ptr, d["state"].ptr, d["mat"].ptr, d["counts"].ptr, d[cur].ptr,
                     d["counts"].ptr if step else None, d[prev].ptr if step else None,
                     d[iprev].ptr if step else None, int(step > 0), d[icur].ptr, d["vc"].ptr) == 0
            ctx.synchronize()
            new = oracle.collide(pool, shapes, pairs, steps[step], caches=co).reshape(n, 1)
            oi, ov = oracle.contact_init(1, w["bodies"], w["state"], w["material"], np.ones(n, np.uint32), new,
                                         np.ones(n, np.uint32), new if o_man is None else o_man, o_imp,
                                         warm=step > 0)
            gi = d[icur].download(np.zeros((n, 1), IMPULSES_DTYPE))
            gv = d["vc"].download(np.zeros((n, 1), VELOCITY_MANIFOLD_DTYPE))
            assert gi.tobytes() == oi.tobytes(), "step %d impulses" % step
            assert gv.tobytes() == ov.tobytes(), "step %d constraints" % step
            o_man, o_imp = new, oi
        assert oi["persisting"].sum() > 500
    finally:
        for b in d.values():
            b.free()
