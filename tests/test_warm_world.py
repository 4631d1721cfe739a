"""CPU tests of the large warm-world workload (a single world of hull stacks in a
container): the contact list is exactly the fat-AABB overlaps, stacked bodies rest at
their drawn SAT separation (checked with the oracle's SAT queries, an independent
computation of the placement), the motion is periodic as documented and the body
rotations are b3QuatMat33 of the orientations."""
import numpy as np
import pytest

from bounce_amd.workloads import (gen_warm_world, quat_xf_rotation, warm_world_pair_inputs, warm_world_state,
                                  warm_world_state_index)
from util import quat_matrix


@pytest.fixture(scope="module")
def world(lib):
    return gen_warm_world(lib, 5, n_hulls=400, columns=5, half_width=4.0)


def _fat_aabbs(world, lib):
    xfs, _ = warm_world_state(world, 0)
    pool = world["pool"]
    lo = np.zeros((len(xfs), 3))
    hi = np.zeros((len(xfs), 3))
    for b in range(len(xfs)):
        v = pool.hull(int(world["body_shapes"][b]))[0].astype(np.float64)
        R = np.stack([xfs[b, 0:3], xfs[b, 3:6], xfs[b, 6:9]], 1).astype(np.float64)
        w = v @ R.T + xfs[b, 9:12]
        lo[b], hi[b] = w.min(0) - 0.2, w.max(0) + 0.2
    return lo, hi


def test_warm_world_contacts_are_the_fat_aabb_overlaps(lib, world):
    c = world["contacts"].astype(np.int64)
    assert (c[:, 0] < c[:, 1]).all()
    key = c[:, 0] * len(world["body_shapes"]) + c[:, 1]
    assert (np.diff(key) > 0).all()  # sorted, unique
    lo, hi = _fat_aabbs(world, lib)
    B = len(lo)
    a, b = np.triu_indices(B, 1)
    ov = np.all((lo[a] <= hi[b]) & (lo[b] <= hi[a]), axis=1)
    ov &= ~((a < world["statics"]) & (b < world["statics"]))  # static pairs never collide
    want = a[ov] * B + b[ov]
    # The generator works from float64 positions, this check from the float32 transforms:
    # pairs whose fat AABBs meet within 1e-5 m may differ.
    gap = np.max(np.maximum(lo[a] - hi[b], lo[b] - hi[a]), axis=1)
    amb = set((a[np.abs(gap) < 1e-5] * B + b[np.abs(gap) < 1e-5]).tolist())
    assert set(key.tolist()) ^ set(want.tolist()) <= amb


def test_warm_world_stacks_rest_at_their_gap(lib, world, oracle):
    """Each body above layer 0 is placed so that its SAT separation from the body below
    is the drawn gap: the oracle's face and edge queries on the step-0 transforms agree."""
    C = world["columns"] ** 2
    ns = world["statics"]
    c = world["contacts"].astype(np.int64)
    vert = np.nonzero((c[:, 0] >= ns) & (c[:, 1] - c[:, 0] == C))[0]
    assert len(vert) == len(world["body_shapes"]) - ns - C
    xfs, sweeps = warm_world_state(world, 0)
    pairs, pxf, _ = warm_world_pair_inputs(world, xfs, sweeps, c[vert])
    q = oracle.queries(world["pool"], pairs, pxf)
    sep = np.maximum(np.maximum(q[:, 1].view(np.float32), q[:, 3].view(np.float32)), q[:, 6].view(np.float32))
    # The drawn gaps: N(-5 mm, 5 mm); separations from float32 transforms within 1e-4 m.
    rng_gap = world["stack_gaps"][c[vert, 1] - ns]
    np.testing.assert_allclose(sep, rng_gap, atol=1e-4)
    assert 0.7 < (sep <= 0).mean() < 0.95


def test_warm_world_motion_is_periodic(world):
    P = world["period"]
    for s in (1, 7, P):
        a, b = warm_world_state(world, s), warm_world_state(world, s + P)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        assert warm_world_state_index(world, s) == warm_world_state_index(world, s + P)
    assert warm_world_state_index(world, 0) == 0
    x0, s0 = warm_world_state(world, 0)
    np.testing.assert_array_equal(s0[:, 0], s0[:, 1])  # first step: orientation0 = orientation
    np.testing.assert_allclose(x0[:, 9:12], world["pos0"], atol=1e-5)
    x1, s1 = warm_world_state(world, 1)
    np.testing.assert_array_equal(s1[:, 0], s0[:, 1])  # orientation0 = previous orientation


def test_quat_xf_rotation_is_the_rotation_of_the_quaternion(world):
    _, sweeps = warm_world_state(world, 3)
    q = sweeps[:, 1]
    R = quat_xf_rotation(q).astype(np.float64).reshape(-1, 3, 3)  # [n][column][row]
    want = quat_matrix(q)
    np.testing.assert_allclose(np.swapaxes(R, 1, 2), want, atol=1e-6)
