"""GPU parity of the step before the narrowphase: the device b3BroadPhase
(b3amd_broadphase_*) and b3Body::SynchronizeShapes' AABBs (b3amd_shape_aabbs) against the
oracle (b3BroadPhase over b3DynamicTree restated, pinned to the reference in
tests/test_broadphase_oracle.py), bit for bit: proxy ids, MoveProxy results, fat AABBs and
every FindPairs list in AddPair order."""
import numpy as np
import pytest

import bounce_amd
import oracle_lib
from util import broadphase_script, run_broadphase_script, ScalarBroadPhase, mixed_workload

pytestmark = pytest.mark.gpu


def _oracle_bp():
    return oracle_lib.BroadPhase(oracle_lib.oracle().lib, "b3o_bp_")


@pytest.fixture(scope="module")
def ctx():
    with bounce_amd.Context(0) as c:
        yield c


@pytest.mark.parametrize("seed,n0,steps", [(1, 250, 15), (2, 600, 10), (3, 20000, 4)])
def test_gpu_broadphase_equals_oracle(ctx, seed, n0, steps):
    print("broadphase script seed %d, %d proxies, %d steps" % (seed, n0, steps), flush=True)
    ops = broadphase_script(seed, n0=n0, steps=steps, extent=20.0 if n0 < 5000 else 120.0)
    ref = _oracle_bp()
    with ctx.broadphase() as bp:
        ida, mva, fa = run_broadphase_script(bp, ops)
        idb, mvb, fb = run_broadphase_script(ScalarBroadPhase(ref), ops)
        assert np.array_equal(ida, idb)
        for x, y in zip(mva, mvb):
            assert np.array_equal(x, np.asarray(y, np.uint32))
        assert len(fa) == len(fb)
        for k, (x, y) in enumerate(zip(fa, fb)):
            assert np.array_equal(x, y), "find %d: %d vs %d pairs" % (k, len(x), len(y))
        assert sum(len(x) for x in fa) > 500
        live = set(range(len(ida)))
        for op in ops:
            if op[0] == "destroy":
                live -= set(op[1].tolist())
        count, cap = bp.info()
        assert count == len(live)
        fat, users, flags = bp.read(0, cap)
        hs = np.array(sorted(live))
        want = np.stack([ref.aabb(idb[h]) for h in hs])
        assert np.array_equal(fat[ida[hs]].view(np.uint32), want.view(np.uint32))
        assert np.array_equal(users[ida[hs]], hs.astype(np.uint32))
        assert (flags[ida[hs]] & 1).all() and int((flags & 1).sum()) == len(live)
        rng = np.random.default_rng(seed)
        pairs = rng.choice(hs, (500, 2))
        pairs = pairs[pairs[:, 0] != pairs[:, 1]]
        got = bp.test_overlap(ida[pairs])
        assert np.array_equal(got, [ref.test_overlap(idb[a], idb[b]) for a, b in pairs])


def test_gpu_broadphase_edge_cases(ctx):
    ref = _oracle_bp()
    box = np.array([[0, 0, 0, 1, 1, 1]], np.float32)
    with ctx.broadphase() as bp:
        assert len(bp.find_pairs()) == 0                       # no proxies
        a = bp.create(box, [7])
        assert a[0] == ref.create(box[0], 7)
        assert len(bp.find_pairs()) == 0 and len(ref.find_pairs()) == 0  # one proxy: nothing to pair
        b = bp.create(box + 0.5, [8])
        assert b[0] == ref.create(box[0] + 0.5, 8)
        assert np.array_equal(bp.find_pairs(), ref.find_pairs())
        assert len(bp.find_pairs()) == 0                       # nothing moved since
        bp.destroy(a), ref.destroy(int(a[0]))
        bp.destroy(b), ref.destroy(int(b[0]))
        assert bp.info()[0] == 0 and len(bp.find_pairs()) == 0
        # ids come back in the reference's order after everything was destroyed
        c = bp.create(np.concatenate([box, box + 2.0, box + 0.25]), [1, 2, 3])
        assert list(c) == [ref.create(box[0], 1), ref.create(box[0] + 2.0, 2), ref.create(box[0] + 0.25, 3)]
        assert np.array_equal(bp.find_pairs(), ref.find_pairs())
        bp.touch(c[1:2]), ref.touch(int(c[1]))
        assert np.array_equal(bp.find_pairs(), ref.find_pairs())
        # a move that stays inside the fat box reports nothing moved
        moved = bp.move(c[:1], box + 0.05, np.zeros((1, 3), np.float32))
        assert moved[0] == 0 == ref.move(int(c[0]), box[0] + 0.05, np.zeros(3, np.float32))
        with pytest.raises(bounce_amd.BounceError):
            bp.destroy([999])
        bp.destroy(c[2:3])
        with pytest.raises(bounce_amd.BounceError):
            bp.touch(c[2:3])  # destroyed
        # a device move naming a proxy that is not live is reported by the next find_pairs
        bp.move(c[2:3], box, np.zeros((1, 3), np.float32))
        with pytest.raises(bounce_amd.BounceError):
            bp.find_pairs()
        assert len(bp.find_pairs()) == 0


def test_gpu_broadphase_query_overflowing_its_stack(ctx):
    """One moved proxy whose box covers 4,000 others: the query's group of lanes meets more
    pending nodes than its LDS stack holds, and the overflow is walked lane by lane."""
    print("broadphase query over 4,000 covered proxies", flush=True)
    ref = _oracle_bp()
    rng = np.random.default_rng(9)
    lo = rng.uniform(0.0, 20.0, (4000, 3)).astype(np.float32)
    boxes = np.concatenate([lo, lo + 0.1], axis=1).astype(np.float32)
    big = np.array([[-1, -1, -1, 21, 21, 21]], np.float32)
    with ctx.broadphase() as bp:
        ids = bp.create(boxes, np.arange(4000))
        assert list(ids) == [ref.create(b, i) for i, b in enumerate(boxes)]
        assert np.array_equal(bp.find_pairs(), ref.find_pairs())
        c = bp.create(big, [4000])
        assert c[0] == ref.create(big[0], 4000)
        got, want = bp.find_pairs(), ref.find_pairs()
        assert len(got) == 4000
        assert np.array_equal(got, want)


def test_gpu_shape_aabbs_equal_oracle(ctx):
    from bounce_amd.types import SWEEP_DTYPE
    from bounce_amd.workloads import quat_xf_rotation
    lib = bounce_amd.load_library()
    pool, shapes, _, _ = mixed_workload(lib, 31, 16, n_hulls=48)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    rng = np.random.default_rng(3)
    nb = 3000
    q0 = rng.normal(size=(nb, 4))
    q0 /= np.linalg.norm(q0, axis=1, keepdims=True)
    q1 = q0 + rng.normal(scale=0.1, size=(nb, 4))
    q1 /= np.linalg.norm(q1, axis=1, keepdims=True)
    q1[::7] = q0[::7] * 1.5  # unnormalised sweep orientations: GetTransform normalises
    c0 = rng.uniform(-50, 50, (nb, 3))
    c1 = c0 + rng.normal(scale=0.3, size=(nb, 3))
    sw = np.zeros(nb, SWEEP_DTYPE)
    sw["localCenter"] = rng.normal(scale=0.2, size=(nb, 3))
    sw["localCenter"][::3] = 0.0
    sw["orientation0"], sw["worldCenter0"], sw["orientation"], sw["worldCenter"] = q0, c0, q1, c1
    xf = np.zeros((nb, 12), np.float32)
    xf[:, :9] = quat_xf_rotation((q1 / np.linalg.norm(q1, axis=1, keepdims=True)).astype(np.float32))
    xf[:, 9:] = c1
    si = rng.integers(0, len(shapes), nb * 2).astype(np.uint32)
    bo = rng.integers(0, nb, nb * 2).astype(np.uint32)
    for sweeps in (sw, None):
        got_a, got_d = ctx.shape_aabbs(si, bo, xf, sweeps)
        want_a, want_d = oracle_lib.shape_aabbs(pool, shapes, si, bo, xf, sweeps)
        assert np.array_equal(got_a.view(np.uint32), want_a.view(np.uint32))
        assert np.array_equal(got_d.view(np.uint32), want_d.view(np.uint32))
    bad = si.copy()
    bad[5] = len(shapes) + 3
    with pytest.raises(bounce_amd.BounceError):
        ctx.shape_aabbs(bad, bo, xf, sw)
