"""GPU parity of the mesh path (device tree query, per-triangle narrowphase, clustering)
against the oracle and the compiled reference's fixture, bit-exact."""
import numpy as np
import pytest

import oracle_lib
from util import assert_manifolds_equal, load_golden, mesh_workload

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
