"""The SAT kernel's edge query tests the bounding caps of two Gauss-map arcs on packed
halves before the exact b3IsMinkowskiFace test (sat.cpp:60-71): five half fused
multiply-adds of s1 s2 - c1 c2 + M1.M2 + 2^-7 and the sign of the result.  The kernel
relies on two facts, checked here on the CPU by emulating that arithmetic (each half
fused multiply-add is exact in float64 and rounded once to float16):

* rounding the inputs to halves and the five roundings change the left-hand side by at
  most 14 * 2^-11 < 2^-7, so the test with 2^-7 added accepts every pair of caps that
  meet;
* the caps (radius half the arc's angle + 0.01 rad) contain the arcs, so every edge pair
  that passes the exact test in FP32 passes the cap test.

A plausible mistake (a dropped term, the slack with the wrong sign, caps built from the
wrong face pair) makes the second check reject exact hits."""
import numpy as np

from bounce_amd.workloads import gen_hull_pairs, gen_hull_pool

SLACK = 2.0 ** -7
U16 = 2.0 ** -11


def _h16(x):
    return np.asarray(x, np.float64).astype(np.float16).astype(np.float64)


def _fma16(a, b, c):
    return _h16(a * b + c)


def _arc_caps(U, V):
    """The host's arc_cap: centre (U + V)/|U + V| and cos/sin of atan2(|U - V|, |U + V|) + 0.01,
    in float64, stored as float32."""
    u = U / np.linalg.norm(U, axis=1, keepdims=True)
    v = V / np.linalg.norm(V, axis=1, keepdims=True)
    s, d = u + v, u - v
    ls, ld = np.linalg.norm(s, axis=1), np.linalg.norm(d, axis=1)
    a = np.arctan2(ld, ls) + 0.01
    full = (ls < 1e-9) | (a >= np.pi)
    m = np.where(full[:, None], 0.0, s / np.where(full, 1.0, ls)[:, None])
    return (m.astype(np.float32), np.where(full, -1.0, np.cos(a)).astype(np.float32),
            np.where(full, 0.0, np.sin(a)).astype(np.float32))


def _edges(pool, h):
    v, e, f, p, c = pool.hull(h)
    i = np.arange(0, len(e), 2)
    U, V = p[e[i, 2], :3], p[e[i + 1, 2], :3]
    return U.astype(np.float32), V.astype(np.float32), (v[e[i + 1, 0]] - v[e[i, 0]]).astype(np.float32)


def test_half_cap_filter_is_conservative(lib):
    pool = gen_hull_pool(lib, 77, 512, 8, 64, 0.25, 2.0)
    pairs, xfs = gen_hull_pairs(lib, 78, pool, 160, -0.0127, 0.05)
    tests = hits = passed = 0
    worst = 0.0
    for (a, b), x in zip(pairs, xfs):
        R1 = np.stack([x[0, 0:3], x[0, 3:6], x[0, 6:9]], 1).astype(np.float64)
        R2 = np.stack([x[1, 0:3], x[1, 3:6], x[1, 6:9]], 1).astype(np.float64)
        R = (R2.T @ R1).astype(np.float32)  # hull 1 into hull 2's frame
        U1, V1, E1 = _edges(pool, a)
        U2, V2, E2 = _edges(pool, b)
        U1, V1, E1 = U1 @ R.T, V1 @ R.T, E1 @ R.T
        # b3IsMinkowskiFace(U1, V1, -E1, -U2, -V2, -E2) in FP32.
        A, B, nE = U1[:, None], V1[:, None], -E1[:, None]
        C, D, DxC = -U2[None], -V2[None], -E2[None]
        dot = lambda p, q: (p[..., 0] * q[..., 0] + p[..., 1] * q[..., 1]) + p[..., 2] * q[..., 2]
        cba, dba, adc, bdc = dot(C, nE), dot(D, nE), dot(A, DxC), dot(B, DxC)
        exact = (cba * dba < 0) & (adc * bdc < 0) & (cba * bdc > 0)
        # Caps: hull 1's arc in hull 2's frame, hull 2's negated arc (the record's -m).
        m1, c1, s1 = _arc_caps(U1, V1)
        m2, c2, s2 = _arc_caps(-U2, -V2)
        g64 = SLACK + s1[:, None].astype(np.float64) * s2[None] - c1[:, None].astype(np.float64) * c2[None] + \
            m1.astype(np.float64) @ m2.T.astype(np.float64)
        g = _fma16(_h16(s1)[:, None], _h16(s2)[None], SLACK)
        g = _fma16(_h16(-c1)[:, None], _h16(c2)[None], g)
        for k in range(3):
            g = _fma16(_h16(m1[:, k])[:, None], _h16(m2[:, k])[None], g)
        ok = g >= 0.0
        worst = max(worst, float(np.abs(g - g64).max()))
        assert not (exact & ~ok).any(), "an exact hit failed the half cap test"
        tests += exact.size
        hits += int(exact.sum())
        passed += int(ok.sum())
    assert worst <= 14 * U16, worst
    assert hits > 1000 and 0.004 < hits / tests < 0.02
    assert passed / tests < 0.1  # the filter still rejects the bulk
