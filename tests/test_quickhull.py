"""The hull builder (b3amd_quickhull, csrc/quickhull.cpp) restates b3QHull::Set
(qhull.cpp:104-354) over qhHull (qh_hull.cpp): the vertex, half-edge and face numbering it
produces is the one every feature key of the narrowphase refers to.

Pinned three ways: record for record against the reference's own b3QHull::Set compiled from
its sources (where oracle/_ref is built), against golden digests the reference wrote
(tests/golden/qhull_digests.npz, so the check also runs where the reference is absent), and
by what any correct hull must satisfy whatever its numbering: Euler's formula, twins
adjacent, closed face loops, unit planes through their face's vertices, every input point
inside, and the centroid inside.  The reference leaves one record undefined: a vertex that
quickhull keeps in its list but no face uses is never written (qhull.cpp:317-320), and the
centroid averages it in; the comparison covers every other record (util.qhull_defined)."""
import numpy as np
import pytest

from util import qhull_clouds, qhull_defined, qhull_digest, run_qhull


@pytest.mark.parametrize("simplify", [0, 1])
def test_quickhull_equals_reference(lib, reference, simplify):
    clouds = qhull_clouds(2028)
    orphans = 0
    for k, p in enumerate(clouds):
        got, go = qhull_defined(run_qhull(lib.b3amd_quickhull, p, simplify))
        want, wo = qhull_defined(run_qhull(reference.lib.b3ref_qhull, p, simplify, returns_status=False))
        assert go == wo, k
        orphans += go
        for name, a, b in zip(("counts", "vertices", "edges", "faces", "planes", "centroid"), got, want):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), "cloud %d (%d points) %s" % (k, len(p), name)
    assert orphans < 10


def test_quickhull_equals_golden_digests(lib):
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "qhull_digests.npz"))
    clouds = np.split(g["points"], np.cumsum(g["sizes"])[:-1])
    assert len(clouds) == len(g["digests"]) > 200
    for k, p in enumerate(clouds):
        for s in (0, 1):
            assert bytes.fromhex(qhull_digest(run_qhull(lib.b3amd_quickhull, p, s))) == g["digests"][k, s].tobytes(), \
                "cloud %d simplify %d" % (k, s)


@pytest.mark.parametrize("simplify", [0, 1])
def test_quickhull_hulls_are_valid(lib, simplify):
    clouds = qhull_clouds(2029, n_ellipsoid=200, n_box=40, n_slab=0)
    for k, p in enumerate(clouds):
        counts, V, E, F, P, C = run_qhull(lib.b3amd_quickhull, p, simplify)
        if counts[0] == 0:
            # left empty: the degenerate special sets (coplanar quad), and with simplification
            # clouds whose dual keeps fewer than 4 directions (qhull.cpp:211-215)
            assert k >= 240 or simplify, k
            continue
        nv, nhe, nf = (int(x) for x in counts)
        assert nv - nhe // 2 + nf == 2, k
        assert np.array_equal(E[:, 1], np.arange(nhe) ^ 1)
        seen = np.zeros(nhe, bool)
        for fi in range(nf):
            h0 = h = F[fi]
            while True:
                assert E[h, 2] == fi and not seen[h]
                seen[h] = True
                # the plane passes through the face's vertices
                assert abs(float(P[fi, :3].astype(np.float64) @ V[E[h, 0]] - P[fi, 3])) < 1e-4
                h = E[h, 3]
                if h == h0:
                    break
        assert seen.all()
        n = P[:, :3].astype(np.float64)
        np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)
        if not simplify:  # the simplified hull merges faces within ~45 degrees and cuts corners
            assert (p.astype(np.float64) @ n.T - P[:, 3]).max() < 1e-4, k
        assert (V.astype(np.float64) @ n.T - P[:, 3]).max() < 1e-4, k  # convex
        assert (C.astype(np.float64) @ n.T - P[:, 3]).max() < 0, k
        if not simplify and k < 200:
            # points on an ellipsoid are all in convex position: every one is a vertex
            assert nv == len(p), k
