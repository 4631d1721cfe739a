"""The step after the narrowphase in the oracle: warm-start matching (b3Manifold::Initialize(old)
in b3Contact::Update's loop, contact.cpp:63-91, manifold.cpp:21-55, with the tangent and
motor impulses taken from the last old manifold, T21) and
b3ContactSolver::InitializeConstraints (contact_solver.cpp:70-276: the world manifold,
manifold.cpp:57-95, normal and tangent effective masses, restitution bias, twist mass).

Pinned bit-exact against the compiled reference's own b3Manifold and b3ContactSolver, for one
and three manifold slots per contact, with and without warm starting; and against
mathematics: every new point persists exactly when an old manifold holds its key and takes
the impulse of the last such manifold, normalMass inverts the effective mass computed in
FP64, and the tangent mass matrix inverts the 2x2 friction block."""
import numpy as np
import pytest

from util import contact_init_workload


def _run(checker, w, M, warm):
    return checker.contact_init(M, w["bodies"], w["state"], w["material"], w["counts"], w["manifolds"],
                                w["old_counts"], w["old_manifolds"], w["old_impulses"], warm)


@pytest.mark.parametrize("M,warm", [(1, True), (3, True), (1, False)])
def test_contact_init_equals_reference(lib, oracle, reference, M, warm):
    w = contact_init_workload(lib, oracle, 41 + M, 1500, M)
    io, vo = _run(oracle, w, M, warm)
    ir, vr = _run(reference, w, M, warm)
    for j in range(M):
        live = w["counts"] > j
        assert io[live, j].tobytes() == ir[live, j].tobytes(), "impulses, slot %d" % j
        assert vo[live, j].tobytes() == vr[live, j].tobytes(), "velocity constraints, slot %d" % j
    touching = vo["pointCount"][w["counts"] > 0, 0] > 0
    assert touching.mean() > 0.2
    if warm:
        assert io["persisting"].sum() > 300
        assert (vo["points"]["velocityBias"] > 0).sum() > 50  # the restitution branch ran


def test_contact_init_math(lib, oracle):
    M = 3
    w = contact_init_workload(lib, oracle, 77, 800, M)
    imp, vc = _run(oracle, w, M, True)
    # Warm-start matching: the impulse of the last old manifold holding the key.
    def keys(m, k):
        p = m["points"][k]
        return (int(p["triangleKey"]),) + tuple(int(x) for x in p["key"])
    for i in range(len(w["counts"])):
        for j in range(int(w["counts"][i])):
            m = w["manifolds"][i, j]
            for k in range(int(m["pointCount"])):
                want_p, want_i = 0, 0.0
                for o in range(int(w["old_counts"][i])):
                    om = w["old_manifolds"][i, o]
                    for q in range(int(om["pointCount"])):
                        if keys(om, q) == keys(m, k):
                            want_p, want_i = 1, w["old_impulses"][i, o]["normalImpulse"][q]
                            break
                assert imp[i, j]["persisting"][k] == want_p and imp[i, j]["normalImpulse"][k] == np.float32(want_i)
            if w["old_counts"][i]:
                last = w["old_impulses"][i, int(w["old_counts"][i]) - 1]
                assert imp[i, j]["motorImpulse"] == last["motorImpulse"]
    # Effective masses in FP64.
    st = w["state"]
    checked = 0
    for i in range(len(w["counts"])):
        a, b = (int(x) for x in w["bodies"][i])
        mA, mB = float(st["invMass"][a]), float(st["invMass"][b])
        iA = st["invI"][a].reshape(3, 3).T.astype(np.float64)
        iB = st["invI"][b].reshape(3, 3).T.astype(np.float64)
        for j in range(int(w["counts"][i])):
            v = vc[i, j]
            for k in range(int(v["pointCount"])):
                p = v["points"][k]
                n = p["normal"].astype(np.float64)
                rnA, rnB = np.cross(p["rA"], n), np.cross(p["rB"], n)
                K = mA + mB + rnA @ iA @ rnA + rnB @ iB @ rnB
                if K > 1e-3:
                    assert abs(p["normalMass"] * K - 1.0) < 1e-4
                    checked += 1
            if v["pointCount"]:
                t1, t2 = v["tangent1"].astype(np.float64), v["tangent2"].astype(np.float64)
                assert abs(t1 @ v["normal"]) < 1e-5 and abs(t2 @ v["normal"]) < 1e-5
                r1A, r1B = np.cross(v["rA"], t1), np.cross(v["rB"], t1)
                r2A, r2B = np.cross(v["rA"], t2), np.cross(v["rB"], t2)
                Kt = np.array([[mA + mB + r1A @ iA @ r1A + r1B @ iB @ r1B, r1A @ iA @ r2A + r1B @ iB @ r2B],
                               [0.0, mA + mB + r2A @ iA @ r2A + r2B @ iB @ r2B]])
                Kt[1, 0] = Kt[0, 1]
                T = v["tangentMass"].reshape(2, 2).T.astype(np.float64)
                if abs(np.linalg.det(Kt)) > 1e-3:
                    np.testing.assert_allclose(T @ Kt, np.eye(2), atol=2e-3)
    assert checked > 300
