"""The SAT kernel's edge query tests the bounding caps of two Gauss-map arcs on the
matrix cores before the exact b3IsMinkowskiFace test (sat.cpp:60-71): a half-precision
v_mfma_f32_32x32x8_f16 of the rows (M1, -cos a1, sin a1, 2^-8) against the columns
(-m2, cos a2, sin a2, 1), i.e. s1 s2 - c1 c2 + M1.M2 + 2^-8, and the sign of the result.
The kernel relies on two facts, checked here on the CPU by emulating that arithmetic
(inputs rounded to float16, with and without flushing half denormals, products exact,
the sum rounded to float32) on the caps the library itself stores at upload
(b3amd_hull_edge_caps):

* rounding the inputs to halves changes each product by at most 2^-10 of its magnitude
  and the products' magnitudes sum to at most 3, flushed half denormals (below 2^-14)
  move each of the five terms by less than 2^-14, and the float32 sum adds under 2^-19:
  together under 3.3e-3 < 2^-8, so the test with 2^-8 added accepts every pair of caps
  that meet;
* the caps contain the region the reference's test can accept, so every edge pair that
  passes the exact test in FP32 and that b3Project scores (both edges at least
  LINEAR_SLOP long, |E1 x E2| >= 0.005 |E1| |E2|, sat.cpp:75-95; the others return
  -FLT_MAX and never become the maximum) passes the cap test.  The caps are the shorter
  arc's, widened per edge by 0.001 rad plus 210 times a bound on how far the test's
  circle (normal -E) can stray from the arc's; edges where that bound is large, or whose
  arc is a half circle, get the whole sphere.

The second check runs over a sweep of the headline hull pool and over degenerate hulls
where a plausible mistake in the caps shows: flat two-face plates (normals exactly
opposite, every arc a half circle), knife edges (normals nearly opposite), boxes at
identity and at 45 degrees (exactly perpendicular normals), identical hulls at identity,
and prisms with 12-gon caps.  A cap of the whole sphere encoded as cos = -1 (instead of
the library's m = 0, cos 0, sin 1), a dropped term, the slack with the wrong sign or
caps built from the wrong face pair each reject exact hits here."""
import ctypes

import numpy as np
import pytest

from bounce_amd.shapes import box_hull, identity_xf, polyhedron
from bounce_amd.types import HullPool, ptr
from bounce_amd.workloads import gen_hull_pairs, gen_hull_pool

SLACK = 2.0 ** -8
BOUND = 3 * 2.0 ** -10 * (1 + 2.0 ** -12) + 5 * 2.0 ** -14 + 2.0 ** -19
LINEAR_SLOP = np.float32(0.005)


def _h16(x):
    return np.asarray(x, np.float64).astype(np.float16).astype(np.float64)


def _h16_ftz(x):
    h = _h16(x)
    return np.where(np.abs(h) < 2.0 ** -14, 0.0, h)


def _mfma16(rows, cols, rnd):
    """sum_k rnd(rows[:, k]) * rnd(cols[:, k]) over the eight terms, float64 products and
    sum, rounded to float32."""
    return (rnd(rows) @ rnd(cols).T).astype(np.float32).astype(np.float64)


def lib_caps(lib, pool, h):
    """(m [E,3], cos [E], sin [E]) float32 as the library stores them for hull h."""
    E = int(pool.desc[h, 3]) // 2
    out = np.zeros((E, 5), np.float32)
    f = lib.b3amd_hull_edge_caps
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    ps = pool.struct()
    assert f(ctypes.byref(ps), h, ptr(out)) == 0
    return out[:, :3], out[:, 3], out[:, 4]


def _edges(pool, h):
    v, e, f, p, c = pool.hull(h)
    i = np.arange(0, len(e), 2)
    U, V = p[e[i, 2], :3], p[e[i + 1, 2], :3]
    P, Q = v[e[i, 0]], v[e[i + 1, 0]]
    return U.astype(np.float32), V.astype(np.float32), P.astype(np.float32), Q.astype(np.float32)


def _rot(x):
    return np.stack([x[0:3], x[3:6], x[6:9]], 1).astype(np.float64)


def _dot(p, q):
    return (p[..., 0] * q[..., 0] + p[..., 1] * q[..., 1]) + p[..., 2] * q[..., 2]


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)


def check_pair(pool, caps, a, b, xa, xb, mutate=None):
    """Returns (scored exact hits, all exact hits, cap passes, tests, worst |g - g64|) for the
    edge query of hull a (transform xa) against hull b (xb); asserts that no scored exact
    hit is rejected by the matrix-core cap test."""
    R1, R2 = _rot(xa), _rot(xb)
    R = (R2.T @ R1).astype(np.float32)  # hull 1's rotation in hull 2's frame
    t = (R2.T @ (xa[9:12].astype(np.float64) - xb[9:12])).astype(np.float32)
    U1, V1, P1, Q1 = _edges(pool, a)
    U2, V2, P2, Q2 = _edges(pool, b)
    U1, V1 = U1 @ R.T, V1 @ R.T
    P1, Q1 = P1 @ R.T + t, Q1 @ R.T + t
    E1, E2 = (Q1 - P1).astype(np.float32), (Q2 - P2).astype(np.float32)
    # b3IsMinkowskiFace(U1, V1, -E1, -U2, -V2, -E2) in FP32 (sat.cpp:60-71, 155).
    A, B, BxA = U1[:, None], V1[:, None], -E1[:, None]
    C, D, DxC = -U2[None], -V2[None], -E2[None]
    adc, bdc, cba, dba = _dot(A, DxC), _dot(B, DxC), _dot(C, BxA), _dot(D, BxA)
    exact = (cba * dba < 0) & (adc * bdc < 0) & (cba * bdc > 0)
    # b3Project's early -FLT_MAX returns (sat.cpp:75-95).
    L1 = np.sqrt(_dot(E1, E1)).astype(np.float32)
    L2 = np.sqrt(_dot(E2, E2)).astype(np.float32)
    X = _cross(E1[:, None], E2[None]).astype(np.float32)
    L = np.sqrt(_dot(X, X)).astype(np.float32)
    scored = exact & (L1[:, None] >= LINEAR_SLOP) & (L2[None] >= LINEAR_SLOP) & \
        ~(L < np.float32(0.005) * L1[:, None] * L2[None])
    # The kernel's filter: hull 1's cap centre rotated into hull 2's frame; hull 2's stored
    # negated centre (-m) and cos, sin.
    m1, c1, s1 = caps[a]
    m2, c2, s2 = caps[b]
    M1 = (m1 @ R.T).astype(np.float32)
    M2 = -m2
    g64 = SLACK + s1[:, None].astype(np.float64) * s2[None] - c1[:, None].astype(np.float64) * c2[None] + \
        M1.astype(np.float64) @ M2.T.astype(np.float64)
    na, nb = len(c1), len(c2)
    rows = np.zeros((na, 8))
    cols = np.zeros((nb, 8))
    rows[:, 0:3], rows[:, 3], rows[:, 4], rows[:, 5], rows[:, 7] = M1, -c1, s1, SLACK, 1.0
    cols[:, 0:3], cols[:, 3], cols[:, 4], cols[:, 5], cols[:, 6] = M2, c2, s2, 1.0, 1.0
    if mutate == "drop_sin":  # a planted mistake: the s1 s2 term dropped
        rows[:, 4] = 0.0
    elif mutate == "swap_arc":  # a planted mistake: hull 1's centre not negated back
        rows[:, 0:3] = -M1
    g_keep = _mfma16(rows, cols, _h16)
    g_ftz = _mfma16(rows, cols, _h16_ftz)
    ok = (g_keep >= 0.0) & (g_ftz >= 0.0)
    g = g_keep if np.abs(g_keep - g64).max() >= np.abs(g_ftz - g64).max() else g_ftz
    bad = scored & ~ok
    assert not bad.any(), "hulls %d/%d: %d scored exact hits fail the matrix-core cap test" % (a, b, int(bad.sum()))
    return int(scored.sum()), int(exact.sum()), int(ok.sum()), exact.size, float(np.abs(g - g64).max())


def _sweep(lib, pool, pairs, xfs, mutate=None):
    caps = {h: lib_caps(lib, pool, h) for h in np.unique(pairs)}
    tot = np.zeros(4, np.int64)
    worst = 0.0
    for (a, b), x in zip(pairs, xfs):
        s, e, p, n, w = check_pair(pool, caps, int(a), int(b), x[0], x[1], mutate)
        tot += (s, e, p, n)
        worst = max(worst, w)
    assert worst <= BOUND < SLACK, worst
    return tot


def test_cap_filter_is_conservative_on_headline_pool(lib):
    """1,200 random pairs of the headline mix (~12 M edge pairs)."""
    pool = gen_hull_pool(lib, 77, 1024, 8, 64, 0.25, 2.0)
    pairs, xfs = gen_hull_pairs(lib, 78, pool, 1200, -0.0127, 0.05)
    scored, hits, passed, tests = _sweep(lib, pool, pairs, xfs)
    assert hits > 10000 and 0.004 < hits / tests < 0.02
    assert scored > 0.9 * hits
    assert passed / tests < 0.1  # the filter still rejects the bulk


def test_whole_sphere_caps_are_rare_on_headline_pool(lib):
    pool = gen_hull_pool(lib, 77, 256, 8, 64, 0.25, 2.0)
    n = full = 0
    for h in range(pool.count):
        m, c, s = lib_caps(lib, pool, h)
        w = (c == 0) & (s == 1) & (np.abs(m).sum(1) == 0)
        n += len(c)
        full += int(w.sum())
        assert np.all(np.abs(np.linalg.norm(m[~w], axis=1) - 1) < 1e-6)
        assert np.all((c[~w] > -0.0101) & (s[~w] > 0.00099))  # radius in [0.001, pi/2 + 0.01]
    assert full / n < 0.005


def plate(w, d, sides=4):
    """Flat polygonal plate in the y = 0 plane: two faces (normals +y and -y)."""
    ang = 2 * np.pi * (np.arange(sides) + 0.5) / sides
    v = np.stack([w * np.cos(ang), np.zeros(sides), d * np.sin(ang)], 1)
    top = list(range(sides))[::-1]  # counter-clockwise seen from +y
    return polyhedron(v, [top, list(range(sides))])


def knife(length, width, thickness):
    """Triangular prism along z whose edge at x = 0 is a knife edge: the two long faces'
    normals are nearly opposite."""
    v = np.array([[0, 0, -length], [width, thickness, -length], [width, -thickness, -length],
                  [0, 0, length], [width, thickness, length], [width, -thickness, length]], np.float64)
    faces = [[0, 1, 2], [3, 5, 4], [0, 3, 4, 1], [0, 2, 5, 3], [1, 4, 5, 2]]
    return polyhedron(v, faces)


def prism(r, h, sides=12):
    ang = 2 * np.pi * np.arange(sides) / sides
    v = np.concatenate([np.stack([r * np.cos(ang), -h * np.ones(sides), r * np.sin(ang)], 1),
                        np.stack([r * np.cos(ang), h * np.ones(sides), r * np.sin(ang)], 1)])
    bottom = list(range(sides))
    top = list(range(2 * sides - 1, sides - 1, -1))
    sidesf = [[k, sides + k, sides + (k + 1) % sides, (k + 1) % sides] for k in range(sides)]
    return polyhedron(v, [bottom, top] + sidesf)


def degenerate_pool():
    hulls = [plate(0.5, 0.5), plate(0.4, 0.3, 6), knife(0.5, 0.6, 0.002), knife(0.4, 0.5, 1e-4),
             box_hull(0.5, 0.5, 0.5), box_hull(1.0, 0.01, 1.0), prism(0.5, 0.25), box_hull(0.3, 0.3, 0.3)]
    return HullPool.concat(hulls)


def _quat_rot(q):
    q = q / np.linalg.norm(q)
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def degenerate_pairs(seed=5, per_kind=24):
    """Pairs over degenerate_pool(): every hull pair at identity, at exact 45-degree turns
    about each axis, at small tilts of those and at random orientations, with the centres
    a short distance apart (overlapping or just separated)."""
    rng = np.random.default_rng(seed)
    c, s = np.float32(np.cos(np.pi / 4)), np.float32(np.sin(np.pi / 4))
    fixed = [np.eye(3), np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]]), np.array([[1, 0, 0], [0, c, -s], [0, s, c]]),
             np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]), np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]])]
    pairs, xfs = [], []
    for a in range(8):
        for b in range(8):
            for k in range(per_kind):
                if k < len(fixed):
                    Ra, Rb = np.eye(3), fixed[k]
                elif k < 2 * len(fixed):
                    Ra, Rb = np.eye(3), fixed[k - len(fixed)] @ _quat_rot(np.r_[rng.normal(0, 1e-3, 3), 1.0])
                else:
                    Ra, Rb = _quat_rot(rng.normal(size=4)), _quat_rot(rng.normal(size=4))
                off = rng.normal(size=3)
                off *= rng.uniform(0.0, 0.8) / np.linalg.norm(off)
                if k == 0:
                    off[:] = 0.0  # identical placement
                pairs.append((a, b))
                xfs.append([identity_xf((0, 0, 0), Ra), identity_xf(off, Rb)])
    return np.array(pairs, np.uint32), np.array(xfs, np.float32)


def test_cap_filter_is_conservative_on_degenerate_hulls(lib):
    pool = degenerate_pool()
    pairs, xfs = degenerate_pairs()
    scored, hits, passed, tests = _sweep(lib, pool, pairs, xfs)
    assert scored > 2000  # the sweep does reach scored hits on these hulls


@pytest.mark.parametrize("mutate", ["drop_sin", "swap_arc"])
def test_cap_filter_check_catches_planted_mistakes(lib, mutate):
    """The containment check is not vacuous: each planted mistake in the emulated filter
    rejects scored exact hits on the degenerate sweep.  (The slack itself is pinned by the
    rounding bound that every sweep asserts, not by a hit: the caps' own angular margin
    also covers the rounding on these inputs.)"""
    pool = degenerate_pool()
    pairs, xfs = degenerate_pairs()
    with pytest.raises(AssertionError, match="scored exact hits fail"):
        _sweep(lib, pool, pairs, xfs, mutate)


@pytest.mark.parametrize("sides", [4, 6])
def test_flat_plates_get_whole_sphere_caps(lib, sides):
    """Every edge of a two-face plate has opposite face normals: its arc is a half circle
    that only the whole-sphere cap bounds, so tilted plates keep their edge hits."""
    pool = HullPool.concat([plate(0.5, 0.5, sides), plate(0.5, 0.5, sides)])
    m, c, s = lib_caps(lib, pool, 0)
    assert np.all(m == 0) and np.all(c == 0) and np.all(s == 1)
    R = _quat_rot(np.array([0.3, 0.1, 0.2, 1.0]))
    xa, xb = identity_xf(), identity_xf((0.1, 0.0, 0.05), R)
    caps = {0: lib_caps(lib, pool, 0), 1: lib_caps(lib, pool, 1)}
    scored, hits, passed, tests, worst = check_pair(pool, caps, 0, 1, xa, xb)
    assert scored >= 4 and passed == tests


def test_reflex_and_misaligned_edges_get_whole_sphere_caps(lib):
    """An edge whose direction does not follow V x U (here the box's half-edge directions
    reversed, as a reflex edge would have) cannot be bounded by the shorter arc's cap."""
    v, e, f, p, c = box_hull(0.5, 0.5, 0.5)
    m, cs, sn = lib_caps(lib, HullPool.concat([(v, e, f, p, c)]), 0)
    assert np.all(cs > 0)  # a proper box: every cap is a quarter-circle cap
    # Mirror the vertices through the origin but keep topology and planes: every edge
    # direction flips against its Gauss-map arc.
    m2, c2, s2 = lib_caps(lib, HullPool.concat([(-v, e, f, p, c)]), 0)
    assert np.all(m2 == 0) and np.all(c2 == 0) and np.all(s2 == 1)
