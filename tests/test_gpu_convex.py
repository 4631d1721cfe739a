"""GPU parity of the sphere/capsule narrowphase (GJK, closed forms, capsule clipping)
against the oracle and the compiled reference's fixture, bit-exact."""
import numpy as np
import pytest

from bounce_amd.types import SPHERE, CAPSULE, HULL
from util import assert_manifolds_equal, load_golden, mixed_workload, special_convex_cases

pytestmark = pytest.mark.gpu


def run(ctx, pool, shapes, pairs, xfs):
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    return ctx.collide(pairs, xfs)


@pytest.mark.parametrize("seed", [1, 2])
def test_gpu_mixed_shapes_equal_oracle(ctx, lib, oracle, seed):
    print("case: mixed shapes seed=%d" % seed, flush=True)
    pool, shapes, pairs, xfs = mixed_workload(lib, 100 + seed, 6000)
    assert_manifolds_equal(run(ctx, pool, shapes, pairs, xfs), oracle.collide(pool, shapes, pairs, xfs))


def test_gpu_special_convex_cases(ctx, oracle):
    print("case: special convex", flush=True)
    pool, shapes, pairs, xfs = special_convex_cases()
    assert_manifolds_equal(run(ctx, pool, shapes, pairs, xfs), oracle.collide(pool, shapes, pairs, xfs))


def test_gpu_golden_mixed(ctx):
    print("case: golden mixed", flush=True)
    g = load_golden("mixed_pairs.npz")
    assert_manifolds_equal(run(ctx, g["pool"], g["shapes"], g["pairs"], g["xfs"]), g["out"])
