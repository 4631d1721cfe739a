"""CPU checks of the product library: C-ABI surface, generators, hull records."""
import ctypes
import os
import re

import numpy as np
import pytest

import bounce_amd
from bounce_amd.workloads import gen_hull_pool, gen_hull_pairs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^(?:int|uint64_t|const char\*|void)\s+(b3amd_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol(lib):
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), "libbounce_amd.so does not export %s" % n


def test_create_without_gpu_fails_cleanly(lib):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rc = lib.b3amd_create(0, 0, ctypes.byref(h))
    assert rc == bounce_amd.ERR_NO_DEVICE
    assert b"GPU" in lib.b3amd_last_error()


def test_generators_are_deterministic(lib):
    a = gen_hull_pool(lib, 17, 64)
    b = gen_hull_pool(lib, 17, 64)
    for f in ("desc", "centroids", "vertices", "edges", "faces", "planes"):
        assert np.array_equal(getattr(a, f), getattr(b, f))
    pa, xa = gen_hull_pairs(lib, 3, a, 500)
    pb, xb = gen_hull_pairs(lib, 3, a, 500)
    assert np.array_equal(pa, pb) and np.array_equal(xa, xb)


def test_generated_hulls_are_valid_b3hulls(lib):
    pool = gen_hull_pool(lib, 23, 200, 8, 64)
    for h in range(pool.count):
        v, e, f, p, c = pool.hull(h)
        V, HE, F = len(v), len(e), len(f)
        # quickhull keeps every point in convex position; nearly coplanar triangles merge
        assert 4 <= V <= 64 and V - HE // 2 + F == 2 and F <= 2 * V - 4  # Euler
        assert np.array_equal(e[:, 1], np.arange(HE) ^ 1)             # twins adjacent (sat.cpp:123-128)
        assert np.array_equal(e[e[:, 1], 1], np.arange(HE))
        # every face loop closes, its half-edges name the face, every half-edge is on one loop
        seen = np.zeros(HE, bool)
        for fi in range(F):
            h0, h, k = f[fi], f[fi], 0
            while True:
                assert e[h, 2] == fi and not seen[h]
                seen[h] = True
                h, k = e[h, 3], k + 1
                if h == h0:
                    break
            assert k >= 3
        assert seen.all()
        # unit planes, all vertices behind every plane, centroid strictly inside
        n = p[:, :3].astype(np.float64)
        np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)
        d = v.astype(np.float64) @ n.T - p[:, 3]
        assert d.max() < 1e-5
        assert (c.astype(np.float64) @ n.T - p[:, 3]).max() < 0


def test_workload_overlap_fraction(lib, oracle):
    """The batched hull-hull placement sets the SAT signed gap to g ~ N(-0.0127, 0.05):
    P(g < 0) = 0.6 is the overlapping fraction."""
    pool = gen_hull_pool(lib, 31, 256)
    pairs, xfs = gen_hull_pairs(lib, 32, pool, 4000, -0.0127, 0.05)
    m = oracle.collide(pool, pool.shapes(), pairs, xfs)
    assert abs((m["pointCount"] > 0).mean() - 0.6) < 0.03
