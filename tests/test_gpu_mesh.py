"""GPU parity of the mesh path (device tree query, per-triangle narrowphase, clustering)
against the oracle and the compiled reference's fixture, bit-exact."""
import numpy as np
import pytest

import oracle_lib
from util import assert_manifolds_equal, load_golden, mesh_workload, plate_workload

pytestmark = pytest.mark.gpu


def test_gpu_mesh_golden(ctx):
    print("case: golden mesh contacts", flush=True)
    g = load_golden("mesh_contacts.npz")
    ctx.upload_hulls(g["pool"])
    ctx.upload_shapes(g["shapes"])
    mid = ctx.mesh_id = ctx.upload_mesh(g["mesh_vertices"], g["mesh_triangles"])
    assert np.array_equal(ctx.mesh_tree(mid).view(np.uint32), g["mesh_nodes"].view(np.uint32))
    counts, lists, out, mcount = ctx.collide_mesh(mid, g["shape_a"], g["xfs"])
    np.testing.assert_array_equal(counts, g["tri_counts"])
    for i in range(len(counts)):
        k = min(int(counts[i]), ctx.MESH_TRIANGLE_CAP)
        np.testing.assert_array_equal(lists[i, :k], g["tri_lists"][i, :k])
    np.testing.assert_array_equal(mcount, g["manifold_counts"])
    assert_manifolds_equal(out.reshape(-1), g["out"].reshape(-1), "golden mesh manifolds")


@pytest.mark.parametrize("seed,size", [(1, 32), (2, 96)])
def test_gpu_mesh_equals_oracle(ctx, lib, oracle, seed, size):
    print("case: mesh seed=%d grid=%d" % (seed, size), flush=True)
    pool, shapes, v, t, xfs = mesh_workload(lib, 200 + seed, size, size, 800)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    mid = ctx.upload_mesh(v, t)
    nodes = ctx.mesh_tree(mid)
    sa = np.arange(len(shapes), dtype=np.uint32)
    counts, lists, out, mcount = ctx.collide_mesh(mid, sa, xfs)
    co = oracle_lib.collide_mesh(oracle.lib, "b3o_", pool, shapes, v, t, nodes, sa, xfs)
    np.testing.assert_array_equal(counts, co[0])
    np.testing.assert_array_equal(mcount, co[3])
    assert_manifolds_equal(out.reshape(-1), co[2].reshape(-1), "mesh manifolds")


@pytest.mark.parametrize("seed", [5, 6])
def test_gpu_mesh_large_clusters_equal_oracle(ctx, oracle, seed):
    """Flat boxes on a flat grid: each contact's one cluster gathers 40-100 points from up to
    50 triangle face contacts, beyond the reference's 32-point inline polygon; the reduction
    runs over all of them (contact_cluster.cpp:497-587)."""
    print("case: plates on a flat grid, seed=%d" % seed, flush=True)
    pool, shapes, v, t, xfs = plate_workload(seed, 40)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    mid = ctx.upload_mesh(v, t)
    nodes = ctx.mesh_tree(mid)
    sa = np.arange(len(shapes), dtype=np.uint32)
    counts, lists, out, mcount = ctx.collide_mesh(mid, sa, xfs)
    assert (counts <= ctx.MESH_TRIANGLE_CAP).all()
    co = oracle_lib.collide_mesh(oracle.lib, "b3o_", pool, shapes, v, t, nodes, sa, xfs)
    np.testing.assert_array_equal(counts, co[0])
    np.testing.assert_array_equal(mcount, co[3])
    assert_manifolds_equal(out.reshape(-1), co[2].reshape(-1), "plate manifolds")


def test_gpu_mesh_wide_polygons_equal_oracle(ctx, oracle):
    """40- and 60-gon prisms lying on large flat triangles: cap and triangle clip to polygons
    of 40-63 vertices, beyond the lane narrowphase's 32; the wide pass builds them, equal to
    the oracle bit for bit and unflagged."""
    from util import wide_prism_mesh
    print("case: wide polygons on a mesh", flush=True)
    pool, shapes, v, t, xfs = wide_prism_mesh(4, 60)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    mid = ctx.upload_mesh(v, t)
    nodes = ctx.mesh_tree(mid)
    sa = np.arange(len(shapes), dtype=np.uint32)
    counts, lists, out, mcount = ctx.collide_mesh(mid, sa, xfs)
    co = oracle_lib.collide_mesh(oracle.lib, "b3o_", pool, shapes, v, t, nodes, sa, xfs)
    np.testing.assert_array_equal(counts, co[0])
    np.testing.assert_array_equal(mcount, co[3])
    assert (out["flags"] == 0).all()
    assert_manifolds_equal(out.reshape(-1), co[2].reshape(-1), "wide prisms on a mesh")


def test_gpu_mesh_long_triangle_lists_equal_oracle(ctx, oracle):
    """Fresh contacts over 98-288 triangles (beyond the device lists' 64): the wide-contact
    pass queries them again into its slab, collides every triangle and clusters hundreds of
    points; manifolds equal the oracle's bit for bit, without the overflow flag, and the
    lists' first 64 entries and the counts equal the oracle's too.  Ordinary contacts in the
    same batch take the list path."""
    from util import plate_workload, wide_plate_workload
    print("case: long triangle lists (large plates on a fine grid)", flush=True)
    pool, shapes, v, t, xfs = wide_plate_workload(8, 30)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    mid = ctx.upload_mesh(v, t)
    nodes = ctx.mesh_tree(mid)
    sa = np.arange(len(shapes), dtype=np.uint32)
    counts, lists, out, mcount = ctx.collide_mesh(mid, sa, xfs)
    co = oracle_lib.collide_mesh(oracle.lib, "b3o_", pool, shapes, v, t, nodes, sa, xfs, cap=512)
    assert (co[0] > ctx.MESH_TRIANGLE_CAP).all()
    np.testing.assert_array_equal(counts, co[0])
    np.testing.assert_array_equal(lists, co[1][:, :ctx.MESH_TRIANGLE_CAP])
    np.testing.assert_array_equal(mcount, co[3])
    assert (out["flags"] == 0).all()
    assert_manifolds_equal(out.reshape(-1), co[2].reshape(-1), "long triangle lists")


def test_gpu_mesh_heightfield_full_size(ctx, lib, oracle):
    """The mixed-shapes-on-terrain configuration at full size: a 256 x 256 heightfield
    (131,072 triangles) and 100,000 spheres, capsules and hulls.  All contacts run on the
    device; every 50th is checked against the oracle (tree, triangle lists, clustered
    manifolds, bit-exact); the whole batch keeps the structural invariants."""
    from bounce_amd.workloads import gen_mesh_world
    print("case: heightfield 256x256, 100000 shapes", flush=True)
    pool, shapes, v, t, xfs = gen_mesh_world(lib, 11)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    mid = ctx.upload_mesh(v, t)
    nodes = ctx.mesh_tree(mid)
    np.testing.assert_array_equal(nodes.view(np.uint32),
                                  oracle_lib.build_tree(oracle.lib, "b3o_", v, t).view(np.uint32))
    sa = np.arange(len(shapes), dtype=np.uint32)
    counts, lists, out, mcount = ctx.collide_mesh(mid, sa, xfs)
    assert (counts <= ctx.MESH_TRIANGLE_CAP).all()
    assert (mcount <= 3).all() and (out["pointCount"] <= 4).all()
    assert (out["pointCount"][np.arange(3)[None, :] >= mcount[:, None]] == 0).all()
    assert 0.5 < (mcount > 0).mean() < 0.9
    idx = np.arange(0, len(sa), 50)
    co = oracle_lib.collide_mesh(oracle.lib, "b3o_", pool, shapes, v, t, nodes, sa[idx], xfs[idx])
    np.testing.assert_array_equal(counts[idx], co[0])
    for k, i in enumerate(idx):
        c = int(counts[i])
        np.testing.assert_array_equal(lists[i, :c], co[1][k, :c])
    np.testing.assert_array_equal(mcount[idx], co[3])
    assert_manifolds_equal(out[idx].reshape(-1), co[2].reshape(-1), "heightfield manifolds")


def test_gpu_mesh_device_resident_equals_host_api(ctx, lib):
    """b3amd_collide_mesh_device on resident buffers returns what b3amd_collide_mesh
    returns, on two batches in a row (the second smaller, reusing the scratch), and flags
    a contact whose shape index is out of range instead of reading past the table."""
    from bounce_amd.types import MANIFOLD_DTYPE
    print("case: mesh device-resident entry point", flush=True)
    pool, shapes, v, t, xfs = mesh_workload(lib, 77, 64, 64, 600)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    mid = ctx.upload_mesh(v, t)
    for n in (600, 250):
        sa = np.arange(n, dtype=np.uint32)[::-1].copy()
        x = np.ascontiguousarray(xfs[sa])
        counts, lists, out, mcount = ctx.collide_mesh(mid, sa, x)
        bad = sa.copy()
        bad[7] = len(shapes) + 5
        ds, dx, do, dm = ctx.alloc(sa.nbytes), ctx.alloc(x.nbytes), ctx.alloc(out.nbytes), ctx.alloc(4 * n)
        dc, dl = ctx.alloc(counts.nbytes), ctx.alloc(lists.nbytes)
        ds.upload(bad)
        dx.upload(x)
        ctx.collide_mesh_device(mid, n, ds, dx, do, dm, dc, dl)
        ctx.synchronize()
        got = do.download(np.zeros((n, 3), MANIFOLD_DTYPE))
        gm = dm.download(np.zeros(n, np.uint32))
        keep = np.arange(n) != 7
        assert_manifolds_equal(got[keep].reshape(-1), out[keep].reshape(-1), "device-resident mesh")
        np.testing.assert_array_equal(gm[keep], mcount[keep])
        np.testing.assert_array_equal(dc.download(np.zeros(n, np.uint32))[keep], counts[keep])
        assert gm[7] == 0 and (got[7]["flags"] & 2).all() and (got[7]["pointCount"] == 0).all()
        for b in (ds, dx, do, dm, dc, dl):
            b.free()


def test_gpu_mesh_contacts_persistent_equal_reference(ctx, lib, reference):
    """Mesh contacts kept across eight steps of moving shapes (b3MeshContact in the
    reference, b3amd_mesh_contacts_* here): after creation and after every
    SynchronizeShapes + FindNewPairs the fat AABBs, triangle lists and triangle caches equal
    the reference's; every Collide's clustered manifolds are bit-exact; the statistics
    counters (hull-triangle calls and cache hits, GJK calls and iterations) agree."""
    import bounce_amd
    from util import HULL_CACHE_FIELDS, assert_caches_equal, defined_hull_cache, moving_mesh_steps
    print("case: persistent mesh contacts, 8 steps", flush=True)
    pool, shapes, v, t, xfs = mesh_workload(lib, 91, 48, 48, 300)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    mid = ctx.upload_mesh(v, t)
    sa = np.arange(len(shapes), dtype=np.uint32)
    steps = moving_mesh_steps(xfs, 9, 5)
    ref = oracle_lib.RefMeshContacts(reference.lib, pool, shapes, v, t, sa, steps[0][0])

    def same_state(dev, what):
        dc, dl, dca, df = dev.read()
        rc, rl, rca, rf = ref.read()
        np.testing.assert_array_equal(dc, rc, what)
        np.testing.assert_array_equal(df.view(np.uint32), rf.view(np.uint32), what)
        for i in range(len(dc)):
            k = min(int(dc[i]), 64)
            np.testing.assert_array_equal(dl[i, :k], rl[i, :k], what)
            assert_caches_equal(defined_hull_cache(dca[i, :k]), defined_hull_cache(rca[i, :k]),
                                "%s contact %d" % (what, i))
        return int((dc > 0).sum())

    ctx.set_counting(True)
    ctx.reset_counters()
    reference.reset_counters()
    with bounce_amd.MeshContacts(ctx, mid, sa, steps[0][0]) as dev:
        same_state(dev, "after creation")
        requeried = 0
        for s in range(8):
            x, sw, p = steps[s]
            got, gc = dev.collide(x, sw)
            want, wc = ref.collide(x, sw)
            np.testing.assert_array_equal(gc, wc, "step %d manifold counts" % s)
            assert_manifolds_equal(got.reshape(-1), want.reshape(-1), "step %d manifolds" % s)
            before = dev.read()[1].copy()
            dev.synchronize(steps[s + 1][0], p)
            ref.synchronize(steps[s + 1][0], p)
            same_state(dev, "after synchronize %d" % s)
            requeried += int((dev.read()[1] != before).any(1).sum())
        assert requeried > 10  # the fast shapes left their fat boxes
    got_c = ctx.counters()
    want_c = reference.counters()
    ctx.set_counting(False)
    for k in ("convex_calls", "convex_cache_hits", "gjk_calls", "gjk_iters", "gjk_max_iters"):
        assert got_c[k] == want_c[k], (k, got_c[k], want_c[k])
    assert want_c["convex_cache_hits"] > 0
    ref.close()


def test_gpu_mesh_contacts_device_entry_points_equal_host_forms(ctx, lib):
    """Two sets of the same persistent mesh contacts stepped eight times side by side, one
    through the host-array entry points, one through collide/test_overlap/synchronize
    _device on resident buffers: manifolds, counts, overlap flags and the kept state (lists,
    caches, fat boxes) are identical after every step."""
    import bounce_amd
    from bounce_amd.types import MANIFOLD_DTYPE
    from util import moving_mesh_steps
    print("case: persistent mesh contacts, device entry points", flush=True)
    pool, shapes, v, t, xfs = mesh_workload(lib, 93, 48, 48, 300)
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    mid = ctx.upload_mesh(v, t)
    sa = np.arange(len(shapes), dtype=np.uint32)
    n = len(sa)
    steps = moving_mesh_steps(xfs, 9, 5)
    dx, ds, dc0 = ctx.alloc(96 * n), ctx.alloc(64 * n), ctx.alloc(12 * n)
    do, dm, dov = ctx.alloc(720 * n), ctx.alloc(4 * n), ctx.alloc(4 * n)
    try:
        with bounce_amd.MeshContacts(ctx, mid, sa, steps[0][0]) as host, \
                bounce_amd.MeshContacts(ctx, mid, sa, steps[0][0]) as dev:
            for s in range(8):
                x, sw, p = steps[s]
                want, wc = host.collide(x, sw)
                dx.upload(np.ascontiguousarray(x, np.float32))
                ds.upload(np.ascontiguousarray(sw, np.float32))
                dev.collide_device(dx, do, dm, ds)
                ctx.synchronize()
                got = do.download(np.zeros((n, 3), MANIFOLD_DTYPE))
                gc = dm.download(np.zeros(n, np.uint32))
                np.testing.assert_array_equal(gc, wc, "step %d counts" % s)
                assert_manifolds_equal(got.reshape(-1), want.reshape(-1), "step %d manifolds" % s)
                ov_want = host.test_overlap(x)
                dev.test_overlap_device(dx, dov)
                ctx.synchronize()
                np.testing.assert_array_equal(dov.download(np.zeros(n, np.uint32)), ov_want, "step %d overlap" % s)
                host.synchronize(steps[s + 1][0], p)
                dx.upload(np.ascontiguousarray(steps[s + 1][0], np.float32))
                dc0.upload(np.ascontiguousarray(p, np.float32))
                dev.synchronize_device(dx, dc0)
                ctx.synchronize()
                hc, hl, hca, hf = host.read()
                dc, dl, dca, df = dev.read()
                np.testing.assert_array_equal(hc, dc, "step %d counts" % s)
                np.testing.assert_array_equal(hf.view(np.uint32), df.view(np.uint32), "step %d fat boxes" % s)
                for i in range(n):  # entries past a contact's count are not part of its state
                    k = min(int(hc[i]), ctx.MESH_TRIANGLE_CAP)
                    np.testing.assert_array_equal(hl[i, :k], dl[i, :k], "step %d lists" % s)
                    np.testing.assert_array_equal(hca[i, :k].view(np.uint8), dca[i, :k].view(np.uint8),
                                                  "step %d caches" % s)
    finally:
        for b in (dx, ds, dc0, do, dm, dov):
            b.free()


def test_gpu_mesh_contacts_persistent_golden(ctx):
    """The persistent mesh-contact fixture written by the reference's own b3MeshContact,
    replayed through b3amd_mesh_contacts_*: six steps, bit-exact."""
    import bounce_amd
    from util import check_persistent_mesh_fixture
    print("case: persistent mesh contacts fixture", flush=True)
    g = load_golden("mesh_contacts_persistent.npz")
    ctx.upload_hulls(g["pool"])
    ctx.upload_shapes(g["shapes"])
    mid = ctx.upload_mesh(g["mesh_vertices"], g["mesh_triangles"])
    with bounce_amd.MeshContacts(ctx, mid, g["shape_a"], g["xfs"][0]) as dev:
        check_persistent_mesh_fixture(g, dev)
