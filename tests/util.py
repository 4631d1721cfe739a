"""Shared test helpers: manifold comparison, FP64 checks, workloads."""
import numpy as np

from bounce_amd.types import MANIFOLD_DTYPE


def manifold_bytes(m):
    return np.ascontiguousarray(m).view(np.uint8).reshape(len(m), -1)


def assert_manifolds_equal(got, want, what="manifolds"):
    """Bit-exact: point count, feature keys, manifold order and all 9 floats per point."""
    assert got.dtype == MANIFOLD_DTYPE and want.dtype == MANIFOLD_DTYPE
    assert len(got) == len(want)
    eq = (manifold_bytes(got) == manifold_bytes(want)).all(1)
    if not eq.all():
        bad = np.where(~eq)[0]
        i = int(bad[0])
        raise AssertionError("%s: %d of %d differ; first %d:\n got  %r\n want %r" % (
            what, len(bad), len(eq), i, got[i], want[i]))


def xf_matrix(x):
    """(R [3,3], p [3]) in float64 from 12 floats (columns, position)."""
    x = np.asarray(x, np.float64)
    return np.stack([x[0:3], x[3:6], x[6:9]], 1), x[9:12]


def world_separations(m, xfa, xfb, ra=0.0, rb=0.0):
    """FP64 b3WorldManifoldPoint separations (manifold.cpp:57-69) of one manifold."""
    Ra, pa = xf_matrix(xfa)
    Rb, pb = xf_matrix(xfb)
    out = []
    for i in range(int(m["pointCount"])):
        p = m["points"][i]
        nA = Ra @ p["localNormal1"].astype(np.float64)
        cA = Ra @ p["localPoint1"].astype(np.float64) + pa
        cB = Rb @ p["localPoint2"].astype(np.float64) + pb
        out.append(float((cB - cA) @ nA - ra - rb))
    return out


def sat_separation_fp64(pool, ha, hb, xfa, xfb):
    """Brute-force FP64 separating-axis value for two hulls: the maximum over all face
    normals of both hulls and all edge-pair cross products (oriented away from A's
    centroid) of the projection gap.  Equals the SAT result the narrowphase computes
    (max of the face and edge queries) up to FP32 rounding and the Gauss-map filter."""
    va, ea, fa, pa_, ca = pool.hull(ha)
    vb, eb, fb, pb_, cb = pool.hull(hb)
    Ra, ta = xf_matrix(xfa)
    Rb, tb = xf_matrix(xfb)
    A = va.astype(np.float64) @ Ra.T + ta
    B = vb.astype(np.float64) @ Rb.T + tb
    cA = Ra @ ca.astype(np.float64) + ta
    best = -np.inf
    for n in pa_[:, :3].astype(np.float64) @ Ra.T:
        best = max(best, (B @ n).min() - (A @ n).max())
    for n in pb_[:, :3].astype(np.float64) @ Rb.T:
        best = max(best, (A @ n).min() - (B @ n).max())
    EA = A[ea[1::2, 0]] - A[ea[0::2, 0]]
    EB = B[eb[1::2, 0]] - B[eb[0::2, 0]]
    nA = pa_[:, :3].astype(np.float64) @ Ra.T
    nB = pb_[:, :3].astype(np.float64) @ Rb.T
    UB, VB = nB[eb[0::2, 2]], nB[eb[1::2, 2]]
    for i in range(len(EA)):
        # Gauss-map arcs (U1, V1) and (-U2, -V2) must intersect (sat.cpp:60-71).
        U1, V1 = nA[ea[2 * i, 2]], nA[ea[2 * i + 1, 2]]
        ADC, BDC = -(EB @ U1), -(EB @ V1)
        CBA, DBA = UB @ EA[i], VB @ EA[i]
        mink = (CBA * DBA < 0) & (ADC * BDC < 0) & (CBA * BDC > 0)
        N = np.cross(EA[i], EB)
        L = np.linalg.norm(N, axis=1)
        ok = mink & (L > 0.005 * np.linalg.norm(EA[i]) * np.linalg.norm(EB, axis=1))
        N = N[ok] / L[ok, None]
        s = np.sign((A[ea[2 * i, 0]] - cA) @ N.T)
        s[s == 0] = 1
        N = N * s[:, None]
        gaps = (B @ N.T).min(0) - (A @ N.T).max(0)
        if len(gaps):
            best = max(best, gaps.max())
    return best


def hull_workload(lib, seed, n_hulls, n_pairs, gap_mean=-0.0127, gap_sigma=0.05, nmin=8, nmax=64):
    """Pool, shapes, pairs and transforms of the batched hull-hull workload at a given size."""
    from bounce_amd.workloads import gen_hull_pool, gen_hull_pairs
    pool = gen_hull_pool(lib, seed, n_hulls, nmin, nmax)
    pairs, xfs = gen_hull_pairs(lib, seed + 1, pool, n_pairs, gap_mean, gap_sigma)
    return pool, pool.shapes(), pairs, xfs


def load_golden(name):
    """A fixture written by tests/golden/make_golden.py from the compiled reference."""
    import os
    from bounce_amd.types import HullPool
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name))
    pool = HullPool(d["desc"], d["centroids"], d["vertices"], d["edges"], d["faces"], d["planes"])
    g = {k: d[k] for k in d.files}
    g["pool"] = pool
    return g


def assert_caches_equal(got, want, what="caches"):
    """Field-wise b3ConvexCache equality (the two padding bytes of b3SimplexCache are
    not part of the record's value)."""
    for f in got.dtype.names:
        if not np.array_equal(got[f], want[f]):
            bad = np.where((got[f] != want[f]).reshape(len(got), -1).any(1))[0]
            raise AssertionError("%s: field %s differs in %d records; first %d: got %r want %r" % (
                what, f, len(bad), bad[0], got[bad[0]], want[bad[0]]))
