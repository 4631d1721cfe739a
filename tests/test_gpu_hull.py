"""GPU parity of the hull-hull narrowphase (HIP kernels, through the C ABI) against the
oracle and the golden fixtures written by the compiled reference: bit-exact point
counts, feature keys, manifold order, geometry and feature caches."""
import sys

import numpy as np
import pytest

import bounce_amd
from bounce_amd.shapes import box_hull, identity_xf
from bounce_amd.types import HullPool, MANIFOLD_DTYPE, empty_caches
from util import assert_caches_equal, assert_manifolds_equal, hull_workload, load_golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(ctx, pool, shapes, pairs, xfs, sweeps=None, caches=None):
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    return ctx.collide(pairs, xfs, sweeps=sweeps, caches=caches)


def test_gpu_box_face_contact(ctx, oracle):
    say("case: box face contact")
    pool = HullPool.concat([box_hull(1, 2, 1), box_hull(1, 1, 1)])
    xfs = np.stack([identity_xf((0, 2, 0)), identity_xf((0, 0, 0))])[None]
    pairs = np.array([[0, 1]], np.uint32)
    got = run(ctx, pool, pool.shapes(), pairs, xfs)
    assert got["pointCount"][0] == 4
    assert_manifolds_equal(got, oracle.collide(pool, pool.shapes(), pairs, xfs))


@pytest.mark.parametrize("name", ["hull_pairs.npz", "box_pairs.npz"])
def test_gpu_golden(ctx, name):
    say("case: golden %s" % name)
    g = load_golden(name)
    assert_manifolds_equal(run(ctx, g["pool"], g["shapes"], g["pairs"], g["xfs"]), g["out"], name)


def test_gpu_golden_warm(ctx):
    say("case: golden warm")
    g = load_golden("hull_pairs_warm.npz")
    caches = g["caches_in"].copy()
    out = run(ctx, g["pool"], g["shapes"], g["pairs"], g["xfs"], sweeps=g["sweeps"], caches=caches)
    assert_manifolds_equal(out, g["out"], "warm step")
    assert_caches_equal(caches, g["caches_out"])


@pytest.mark.parametrize("seed,nmin,nmax,gap", [(101, 8, 64, -0.0127), (202, 4, 16, 0.0), (303, 48, 64, -0.05)])
def test_gpu_equals_oracle_random(ctx, lib, oracle, seed, nmin, nmax, gap):
    say("case: random seed=%d n=%d..%d gap=%g" % (seed, nmin, nmax, gap))
    pool, shapes, pairs, xfs = hull_workload(lib, seed, 300, 6000, gap_mean=gap, nmin=nmin, nmax=nmax)
    got = run(ctx, pool, shapes, pairs, xfs)
    assert_manifolds_equal(got, oracle.collide(pool, shapes, pairs, xfs))
    caches = empty_caches(len(pairs))
    want_c = empty_caches(len(pairs))
    got2 = run(ctx, pool, shapes, pairs, xfs, caches=caches)
    oracle.collide(pool, shapes, pairs, xfs, caches=want_c)
    assert_manifolds_equal(got2, got)
    assert_caches_equal(caches, want_c)


def test_gpu_large_hulls(ctx, lib, oracle):
    """Hulls beyond the 64-vertex class run on the large-record kernel."""
    say("case: large hulls 70..128")
    pool, shapes, pairs, xfs = hull_workload(lib, 404, 40, 400, nmin=70, nmax=128)
    assert_manifolds_equal(run(ctx, pool, shapes, pairs, xfs), oracle.collide(pool, shapes, pairs, xfs))


def test_gpu_warm_start_sequence(ctx, lib, oracle):
    """Five steps with persistent caches, moving bodies and sweeps: every cache path."""
    say("case: warm-start sequence")
    pool, shapes, pairs, xfs = hull_workload(lib, 505, 200, 4000)
    rng = np.random.default_rng(5)
    cg, co = empty_caches(len(pairs)), empty_caches(len(pairs))
    x = xfs.copy()
    for step in range(5):
        sweeps = np.zeros((len(pairs), 4, 4), np.float32)
        sweeps[:, :, 3] = 1.0
        tilt = rng.uniform(0, 0.15, size=len(pairs)).astype(np.float32)
        sweeps[:, 3, 1] = tilt
        sweeps[:, 3, 3] = np.sqrt(1 - tilt * tilt)
        got = run(ctx, pool, shapes, pairs, x, sweeps=sweeps, caches=cg)
        want = oracle.collide(pool, shapes, pairs, x, sweeps=sweeps, caches=co)
        assert_manifolds_equal(got, want, "step %d" % step)
        assert_caches_equal(cg, co, "step %d" % step)
        x[:, 1, 9:12] += rng.normal(0, 0.003, size=(len(pairs), 3)).astype(np.float32)


def _corner_down_rotation():
    """Rotation matrix taking the box corner direction (1, 1, 1) / sqrt(3) to (0, -1, 0)."""
    a = np.ones(3) / np.sqrt(3.0)
    b = np.array([0.0, -1.0, 0.0])
    v = np.cross(a, b)
    K = np.array([[0.0, -v[2], v[1]], [v[2], 0.0, -v[0]], [-v[1], v[0], 0.0]])
    return np.eye(3) + K + K @ K / (1.0 + a @ b)


def test_gpu_cached_face_bound_at_threshold(ctx, oracle):
    """Face-cached box pairs whose separation crosses the radius sum in half-micrometre
    steps, the upper box standing on a corner over the lower box's top face so that its
    bounding sphere is tight along the face normal: the cached reads that the bounding
    sphere settles and those that search the vertices both equal the reference's."""
    say("case: cached face read at the bounding-sphere threshold")
    pool = HullPool.concat([box_hull(1, 1, 1), box_hull(1, 1, 1)])
    shapes = pool.shapes()
    r = float(shapes["radius"][0]) + float(shapes["radius"][1])
    R = _corner_down_rotation()
    gaps = r + np.linspace(-1e-4, 3e-4, 801)
    rng = np.random.default_rng(17)
    n = len(gaps)
    xfs = np.zeros((n, 2, 12), np.float32)
    for i, g in enumerate(gaps):
        xfs[i, 0] = identity_xf((0.0, 0.0, 0.0))
        xfs[i, 1] = identity_xf((rng.uniform(-0.3, 0.3), 1.0 + np.sqrt(3.0) + g, rng.uniform(-0.3, 0.3)), R)
    pairs = np.tile(np.array([[0, 1]], np.uint32), (n, 1))
    cg, co = empty_caches(n), empty_caches(n)
    for step in range(3):
        got = run(ctx, pool, shapes, pairs, xfs, caches=cg)
        want = oracle.collide(pool, shapes, pairs, xfs, caches=co)
        assert_manifolds_equal(got, want, "step %d" % step)
        assert_caches_equal(cg, co, "step %d" % step)
        xfs[:, 1, 10] += np.float32(2e-6)
    # Both cached states occur, on hull 1's top face.
    assert (co["state"] == 0).any() and (co["state"] == 1).any()
    assert (co["ftype"] == 1).all() and (co["index1"] == 3).all()


def test_gpu_device_api_equals_host_api(ctx, lib):
    say("case: device-resident API")
    pool, shapes, pairs, xfs = hull_workload(lib, 606, 100, 3000)
    want = run(ctx, pool, shapes, pairs, xfs)
    n = len(pairs)
    dp, dx, do = ctx.alloc(pairs.nbytes), ctx.alloc(xfs.nbytes), ctx.alloc(n * MANIFOLD_DTYPE.itemsize)
    dp.upload(pairs)
    dx.upload(xfs)
    ctx.collide_device(n, dp, dx, do)
    ctx.synchronize()
    got = do.download(np.zeros(n, MANIFOLD_DTYPE))
    for b in (dp, dx, do):
        b.free()
    assert_manifolds_equal(got, want)


def test_gpu_budget_admission(lib):
    say("case: HBM budget admission")
    with bounce_amd.Context(0, hbm_budget_bytes=1 << 20) as c:
        base = c.memory_in_use  # the context's own counter block
        assert base < 4096
        with pytest.raises(bounce_amd.BounceError) as e:
            c.alloc(2 << 20)
        assert e.value.code == bounce_amd.ERR_BUDGET
        b = c.alloc(1 << 19)
        assert c.memory_in_use >= base + (1 << 19)
        b.free()
        assert c.memory_in_use == base


def test_gpu_full_size_properties(ctx, lib, oracle):
    """1,048,576 pairs: ~60% overlapping, manifolds well formed, a strided sample of
    2,048 pairs bit-exact against the oracle, and the batch is deterministic."""
    say("case: full size 1M pairs")
    from bounce_amd.workloads import gen_hull_pool, gen_hull_pairs
    pool = gen_hull_pool(lib, 2024, 65536)
    pairs, xfs = gen_hull_pairs(lib, 2025, pool, 1 << 20, -0.0127, 0.05)
    shapes = pool.shapes()
    got = run(ctx, pool, shapes, pairs, xfs)
    frac = (got["pointCount"] > 0).mean()
    assert 0.57 < frac < 0.63
    assert got["pointCount"].max() <= 4 and (got["flags"] == 0).all()
    idx = np.arange(0, len(pairs), 512)
    assert_manifolds_equal(got[idx], oracle.collide(pool, shapes, pairs[idx], xfs[idx]))
    again = ctx.collide(pairs, xfs)
    assert_manifolds_equal(again, got, "repeat")


@pytest.mark.parametrize("waves,chunk", [(6, 8), (5, 4), (4, 1), (0, 0), (2, 64)])
def test_gpu_sat_launch_shapes_agree(ctx, lib, oracle, waves, chunk):
    """Every launch shape of the persistent SAT kernel (resident waves and register budget,
    tail chunk) gives the oracle's manifolds bit for bit, for hulls of 8-64 points and
    beyond 192 unique edges (several row-block passes and hull-1 records read from HBM)."""
    say("case: SAT launch shape waves=%d chunk=%d" % (waves, chunk))
    try:
        ctx.set_sat_tuning(waves, chunk)
        for seed, nmin, nmax, n in ((707, 8, 64, 3000), (808, 70, 128, 300)):
            pool, shapes, pairs, xfs = hull_workload(lib, seed, 60, n, nmin=nmin, nmax=nmax)
            assert_manifolds_equal(run(ctx, pool, shapes, pairs, xfs), oracle.collide(pool, shapes, pairs, xfs))
    finally:
        ctx.set_sat_tuning(0, 0)


def test_gpu_autotune(ctx, lib, oracle):
    """The autotuner times every launch shape on a device batch and keeps a valid one;
    results after tuning still match the oracle."""
    say("case: autotune")
    pool, shapes, pairs, xfs = hull_workload(lib, 909, 400, 20000)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    n = len(pairs)
    dp, dx, do = ctx.alloc(pairs.nbytes), ctx.alloc(xfs.nbytes), ctx.alloc(n * MANIFOLD_DTYPE.itemsize)
    try:
        dp.upload(pairs)
        dx.upload(xfs)
        (waves, chunk), ms = ctx.autotune(n, dp, dx, do, repeats=1)
        say("autotune: waves=%d chunk=%d %.3f ms" % (waves, chunk, ms))
        assert 1 <= chunk <= 64 and 0 <= waves <= 32 and ms > 0
        idx = np.arange(0, n, 7)
        got = ctx.collide(pairs[idx], xfs[idx])
        assert_manifolds_equal(got, oracle.collide(pool, shapes, pairs[idx], xfs[idx]))
    finally:
        for b in (dp, dx, do):
            b.free()
        ctx.set_sat_tuning(0, 0)


def test_gpu_graph_replay_matches_eager(ctx, lib):
    """A captured batch replayed over changing transforms with persistent caches equals
    the eager batch on every replay (eager work and synchronisations in between); a graph
    refuses to launch after the shape table is replaced."""
    say("case: graph replay")
    pool, shapes, pairs, xfs = hull_workload(lib, 1111, 200, 4000)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    n = len(pairs)
    dp, dx, do = ctx.alloc(pairs.nbytes), ctx.alloc(xfs.nbytes), ctx.alloc(n * MANIFOLD_DTYPE.itemsize)
    dc = ctx.alloc(n * empty_caches(1).itemsize)
    try:
        dp.upload(pairs)
        dc.upload(empty_caches(n))
        cur = np.array(xfs, np.float32).reshape(n, 2, 12)
        dx.upload(cur)
        rng = np.random.default_rng(11)
        eager_caches = empty_caches(n)
        with ctx.capture(n, dp, dx, do, d_caches=dc) as g:
            for step in range(4):
                want = ctx.collide(pairs, cur, caches=eager_caches)
                dx.upload(cur)
                g.launch()
                ctx.synchronize()
                got = do.download(np.zeros(n, MANIFOLD_DTYPE))
                got_c = dc.download(empty_caches(n))
                assert_manifolds_equal(got, want, "replay %d" % step)
                assert_caches_equal(got_c, eager_caches)
                cur = cur.copy()
                cur[:, 1, 9:12] += rng.normal(0.0, 0.01, size=(n, 3)).astype(np.float32)
            ctx.upload_shapes(shapes)
            with pytest.raises(bounce_amd.BounceError) as e:
                g.launch()
            assert e.value.code == bounce_amd.ERR_ARGUMENT
    finally:
        for b in (dp, dx, do, dc):
            b.free()


def prism(n, r, h):
    """Right prism over a regular n-gon (radius r, half-height h): two n-gon caps."""
    from bounce_amd.shapes import polyhedron
    a = 2 * np.pi * np.arange(n) / n
    ring = np.stack([r * np.cos(a), np.zeros(n), r * np.sin(a)], 1)
    v = np.concatenate([ring + (0, -h, 0), ring + (0, h, 0)])
    faces = [list(range(n)), list(range(2 * n - 1, n - 1, -1))]  # outward: -y, +y
    faces += [[i, n + i, n + (i + 1) % n, (i + 1) % n] for i in range(n)]
    return polyhedron(v, faces)


def test_gpu_large_faces_fallback(ctx, oracle):
    """Prisms stacked cap to cap: hexagons rotated against each other clip to polygons of up
    to 12 vertices and 12-gon caps start above the group kernel's 8, so these contacts take
    the lane-per-pair path; everything still equals the oracle bit for bit."""
    say("case: large faces (prism caps)")
    rng = np.random.default_rng(17)
    hulls = [prism(n, r, 0.3) for n, r in ((6, 0.6), (8, 0.5), (12, 0.7), (6, 0.45))]
    pool = HullPool.concat(hulls)
    n = 400
    pairs = rng.integers(0, pool.count, (n, 2)).astype(np.uint32)
    xfs = np.zeros((n, 2, 12), np.float32)
    for i in range(n):
        yaw = rng.uniform(0, 2 * np.pi)
        c, s = np.cos(yaw), np.sin(yaw)
        R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
        tilt = rng.normal(0, 0.002, 3)
        R = R @ np.array([[1, -tilt[2], tilt[1]], [tilt[2], 1, -tilt[0]], [-tilt[1], tilt[0], 1]])
        xfs[i, 0] = identity_xf((0, 0, 0))
        xfs[i, 1] = identity_xf((rng.normal(0, 0.1), 0.6 - rng.uniform(0.0, 0.02), rng.normal(0, 0.1)), R)
    want = oracle.collide(pool, pool.shapes(), pairs, xfs)
    assert (want["pointCount"] > 0).mean() > 0.9
    got = run(ctx, pool, pool.shapes(), pairs, xfs)
    assert_manifolds_equal(got, want, "prism caps")

    # Warm steps with feature caches and sweeps: cached face features on 12-gon caps are
    # rebuilt (collide_hulls_cache.cpp:113-153) with an incident polygon of 12 vertices, beyond
    # the group kernel's 8, so they take the lane-per-pair cache fallback.
    say("case: large faces, warm steps with caches")
    cg, co = empty_caches(n), empty_caches(n)
    x = xfs.copy()
    rebuilt_12gon = 0
    for step in range(3):
        sweeps = np.zeros((n, 4, 4), np.float32)
        sweeps[:, :, 3] = 1.0
        if step > 0:
            yaw = rng.uniform(0.0, 0.04, n).astype(np.float32)
            sweeps[:, 3, 1] = yaw
            sweeps[:, 3, 3] = np.sqrt(1.0 - yaw * yaw)
            both12 = (pairs[:, 0] == 2) & (pairs[:, 1] == 2)
            face_overlap = (co["state"] == 1) & (co["ftype"] != 0)
            rebuilt_12gon += int((both12 & face_overlap).sum())
        got = run(ctx, pool, pool.shapes(), pairs, x, sweeps=sweeps, caches=cg)
        want = oracle.collide(pool, pool.shapes(), pairs, x, sweeps=sweeps, caches=co)
        assert_manifolds_equal(got, want, "prism caps, warm step %d" % step)
        assert_caches_equal(cg, co, "prism caps, warm step %d" % step)
        x[:, 1, 9:12] += rng.normal(0.0, 0.0005, size=(n, 3)).astype(np.float32)
    assert rebuilt_12gon >= 10, rebuilt_12gon


def test_gpu_wide_polygons(ctx, oracle):
    """Clip polygons of 48-120 vertices (24-, 40- and 60-gon prism caps stacked and turned),
    beyond the lane kernels' 32: the wide pass builds them with the polygons in LDS, equal to
    the oracle bit for bit and unflagged, cold and over three warm steps with caches and
    sweeps (the cached face rebuilds of these caps take the wide cache pass)."""
    from util import wide_prism_pairs
    say("case: wide clip polygons (24/40/60-gon caps)")
    rng = np.random.default_rng(23)
    pool, pairs, xfs = wide_prism_pairs(3, 300)
    n = len(pairs)
    want = oracle.collide(pool, pool.shapes(), pairs, xfs)
    got = run(ctx, pool, pool.shapes(), pairs, xfs)
    assert (got["flags"] == 0).all()
    assert_manifolds_equal(got, want, "wide prism caps")
    say("case: wide clip polygons, warm steps with caches")
    cg, co = empty_caches(n), empty_caches(n)
    x = xfs.copy()
    for step in range(4):
        sweeps = np.zeros((n, 4, 4), np.float32)
        sweeps[:, :, 3] = 1.0
        if step > 0:
            yaw = rng.uniform(0.0, 0.04, n).astype(np.float32)
            sweeps[:, 3, 1] = yaw
            sweeps[:, 3, 3] = np.sqrt(1.0 - yaw * yaw)
        got = run(ctx, pool, pool.shapes(), pairs, x, sweeps=sweeps, caches=cg)
        want = oracle.collide(pool, pool.shapes(), pairs, x, sweeps=sweeps, caches=co)
        assert (got["flags"] == 0).all()
        assert_manifolds_equal(got, want, "wide prism caps, warm step %d" % step)
        assert_caches_equal(cg, co, "wide prism caps, warm step %d" % step)
        x[:, 1, 9:12] += rng.normal(0.0, 0.0005, size=(n, 3)).astype(np.float32)
    assert ((co["state"] == 1) & (co["ftype"] != 0)).sum() > 100  # face features cached and rebuilt


def test_gpu_degenerate_hulls(ctx, oracle):
    """Flat two-face plates (edge arcs are half circles), knife edges, boxes at identity and
    at 45 degrees, identical placements and 12-gon prisms, against each other in every
    combination: the SAT kernel's arc-cap filter keeps every edge hit the reference scores,
    so manifolds and caches equal the oracle's bit for bit."""
    from test_cap_filter import degenerate_pairs, degenerate_pool
    say("case: degenerate hulls (plates, knife edges, axis-aligned boxes)")
    pool = degenerate_pool()
    pairs, xfs = degenerate_pairs(seed=9, per_kind=24)
    shapes = pool.shapes()
    want = oracle.collide(pool, shapes, pairs, xfs)
    got = run(ctx, pool, shapes, pairs, xfs)
    assert_manifolds_equal(got, want, "degenerate hulls")
    cg, co = empty_caches(len(pairs)), empty_caches(len(pairs))
    got = run(ctx, pool, shapes, pairs, xfs, caches=cg)
    oracle.collide(pool, shapes, pairs, xfs, caches=co)
    assert_manifolds_equal(got, want, "degenerate hulls, caches")
    assert_caches_equal(cg, co, "degenerate hulls")
    plates = (pairs[:, 0] < 2) & (pairs[:, 1] < 2)
    assert (want["pointCount"][plates] > 0).sum() > 20
    # Edge-feature results occur among the plate pairs (the case whole-sphere caps fix).
    assert ((co["ftype"] == 0) & (co["state"] != 2) & plates).sum() > 0
