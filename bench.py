"""Narrowphase throughput on one MI355X: 1,048,576 batched hull-hull pairs.

One step = one b3amd_collide_device batch over all pairs (feature-cache kernel, SAT
kernel, contact kernel) with inputs resident in HBM and no feature caches: every pair takes
b3CollideHulls without a cache (collide_hulls.cpp:206-289), as b3CollideHullAndHull does
with b3_convexCache off (collide_hulls.cpp:300-316).  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--pairs N] [--pool P] [--no-cpu]
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "narrowphase pairs/s (1M hull pairs); p50/p99 per-world step latency, batched"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFLOPS = 157.3          # MI355X FP32 vector peak (packed, FMA = 2)
VALU_UNFUSED_TFLOPS = 39.3        # 256 CU x 64 lanes x 2.4 GHz, one op per lane per clock


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# The CPU share of a one-GPU box is 16 host cores: worker pools stay within it whatever
# nproc reports, and every cpu_baseline object states nproc, the affinity set and the CPU
# model beside the number of processes used.
CPU_SHARE = 16


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0))
    return {"nproc": os.cpu_count(), "affinity_cores": aff, "cpu_model": model,
            "processes": max(1, min(CPU_SHARE, aff))}


def wall_span(spans):
    """Wall time of work timed in several processes: from the first start to the last end
    (time.perf_counter is the system-wide monotonic clock on Linux)."""
    return max(e for s, e in spans) - min(s for s, e in spans)


def workload(lib, n_pairs, n_pool):
    from bounce_amd.workloads import gen_hull_pool, gen_hull_pairs
    pool = gen_hull_pool(lib, 2024, n_pool, 8, 64, 0.25, 2.0)
    pairs, xfs = gen_hull_pairs(lib, 2025, pool, n_pairs, -0.0127, 0.05)
    return pool, pairs, xfs


# ----------------------------------------------------------------- CPU baseline

_CPU_JOB = {}  # set before forking: children inherit the arrays instead of unpickling them
_RAIN_JOB = {}


def _cpu_worker(chunk):
    j = _CPU_JOB
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    chk = oracle_lib.reference() if j["kind"] == "reference" else oracle_lib.oracle()
    pairs, xfs = j["pairs"][chunk], j["xfs"][chunk]
    t = time.perf_counter()
    m = chk.collide(j["pool"], j["shapes"], pairs, xfs)
    return t, time.perf_counter(), int((m["pointCount"] > 0).sum())


def cpu_baseline(pool, pairs, xfs, sample, shapes=None):
    """The reference narrowphase compiled from its own sources (oracle/_ref), or the
    oracle restatement where that is absent, on a strided sample of the same pairs,
    one process per core (the reference writes unsynchronised global counters).  `shapes`
    defaults to one hull shape per pool hull."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    kind = "reference" if oracle_lib.reference() is not None else "port"
    host = host_cpu()
    cores = host["processes"]
    stride = max(1, len(pairs) // sample)
    idx = np.arange(0, len(pairs), stride)[:sample]
    chunks = np.array_split(idx, cores)
    _CPU_JOB.update(kind=kind, pool=pool, shapes=pool.shapes() if shapes is None else shapes, pairs=pairs, xfs=xfs)
    ctx = mp.get_context("fork")
    with ctx.Pool(cores) as p:
        p.map(_cpu_worker, [c[:64] for c in chunks])  # load the library in every worker
        res = p.map(_cpu_worker, chunks, chunksize=1)
    wall = wall_span([(r[0], r[1]) for r in res])
    cpu_s = sum(r[1] - r[0] for r in res)
    return dict({
        "value": len(idx) / wall,
        "unit": "pairs/s",
        "cores": cores,
        "kind": kind,
        "sample": "%d of the %d pairs (every %d-th), %d processes, %.1f CPU-s, wall %.2f s, %s" % (
            len(idx), len(pairs), stride, cores, cpu_s, wall, os.uname().machine),
        "per_core_pairs_per_s": len(idx) / cpu_s,
    }, **host)


# ----------------------------------------------------------------- algorithmic work

def algorithmic_work(pool, pairs, xfs, stride):
    """FP32 ops and bytes per pair of the SAT stage, from the op-counting oracle on a
    strided sample (scaled to the batch).  Bytes count each record the reference
    control flow reads once: transforms, then hull-1 planes + hull-2 vertices (face query
    1), hull-2 planes + hull-1 vertices (face query 2), whole hull records
    (12V + 16HE + 20F + 24 per hull, the b3Hull arrays) when the edge query runs, plus
    the 32-byte query result."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    oc = oracle_lib.oracle_counting()
    idx = np.arange(0, len(pairs), stride)
    oc.lib.b3o_flop_reset()
    oc.collide(pool, pool.shapes(), pairs[idx], xfs[idx])
    sat_flops = oc.lib.b3o_flop_count_phase(1)
    contact_flops = oc.lib.b3o_flop_count_phase(2)
    q = oracle_lib.oracle().queries(pool, pairs[idx], xfs[idx])
    f1, f2, e = q[:, 1].view(np.float32), q[:, 3].view(np.float32), q[:, 6].view(np.float32)
    d = pool.desc
    a, b = pairs[idx, 0], pairs[idx, 1]
    V1, HE1, F1 = d[a, 1].astype(np.int64), d[a, 3].astype(np.int64), d[a, 5].astype(np.int64)
    V2, HE2, F2 = d[b, 1].astype(np.int64), d[b, 3].astype(np.int64), d[b, 5].astype(np.int64)
    full = (12 * V1 + 16 * HE1 + 20 * F1 + 24) + (12 * V2 + 16 * HE2 + 20 * F2 + 24)
    byts = np.where(f1 > 0, 96 + 16 * F1 + 12 * V2,
                    np.where(f2 > 0, 96 + 16 * F1 + 12 * V2 + 16 * F2 + 12 * V1, 96 + full)) + 32
    scale = len(pairs) / len(idx)
    return {"sat_flops": sat_flops * scale, "contact_flops": contact_flops * scale,
            "sat_bytes": float(byts.sum()) * scale, "sample": len(idx),
            "overlap": float(((f1 <= 0) & (f2 <= 0) & (e <= 0)).mean())}


# ----------------------------------------------------------------- FP64 separation error

def separation_error_fp64(out, xfs, radius_a=0.0, radius_b=0.0):
    """Error of every manifold point's separation: b3WorldManifoldPoint::Initialize
    (manifold.cpp:57-69) evaluated in FP32 in the reference's operation order
    (mat33.h:129-132, transform.h:74-77, vec3.h:158-161) against the same formula in FP64
    from the same FP32 local points, normals and transforms."""
    cnt = out["pointCount"].astype(np.int64)
    mask = np.arange(4)[None, :] < cnt[:, None]
    pts = out["points"][mask]
    pair = np.repeat(np.arange(len(out)), cnt)
    xa, xb = xfs[pair, 0], xfs[pair, 1]
    seps = []
    for dt in (np.float32, np.float64):
        A, B = xa.astype(dt), xb.astype(dt)

        def rot(X, v):
            return (v[:, 0:1] * X[:, 0:3] + v[:, 1:2] * X[:, 3:6]) + v[:, 2:3] * X[:, 6:9]

        nA = rot(A, pts["localNormal1"].astype(dt))
        cA = rot(A, pts["localPoint1"].astype(dt)) + A[:, 9:12]
        cB = rot(B, pts["localPoint2"].astype(dt)) + B[:, 9:12]
        d = cB - cA
        seps.append(((d[:, 0] * nA[:, 0] + d[:, 1] * nA[:, 1]) + d[:, 2] * nA[:, 2]) - dt(radius_a) - dt(radius_b))
    err = np.abs(seps[0].astype(np.float64) - seps[1])
    if len(err) == 0:
        return {"points": 0}
    return {"points": int(len(err)), "median_abs_m": float(np.median(err)),
            "p99_abs_m": float(np.percentile(err, 99)), "max_abs_m": float(err.max()),
            "max_abs_over_abs_separation_ge_1mm": float((err / np.maximum(np.abs(seps[1]), 1e-3)).max())}


# ----------------------------------------------------------------- mixed shapes vs terrain

_MESH_JOB = {}


def _mesh_cpu_worker(chunk):
    j = _MESH_JOB
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    ref = oracle_lib.reference()
    lib, prefix = (ref.lib, "b3ref_") if ref is not None else (oracle_lib.oracle().lib, "b3o_")
    # A first call on a few contacts builds the mesh's tree (b3Mesh::BuildTree), which the
    # timed call then reuses.
    oracle_lib.collide_mesh(lib, prefix, j["pool"], j["shapes"], j["verts"], j["tris"], j["nodes"],
                            chunk[:8].astype(np.uint32), j["xfs"][chunk[:8]])
    t0 = time.perf_counter()
    oracle_lib.collide_mesh(lib, prefix, j["pool"], j["shapes"], j["verts"], j["tris"], j["nodes"],
                            chunk.astype(np.uint32), j["xfs"][chunk])
    return t0, time.perf_counter()


def mesh_cpu_baseline(pool, shapes, verts, tris, xfs, sample):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    kind = "reference" if oracle_lib.reference() is not None else "port"
    nodes = oracle_lib.build_tree(oracle_lib.oracle().lib, "b3o_", verts, tris)
    host = host_cpu()
    cores = host["processes"]
    idx = np.arange(0, len(shapes), max(1, len(shapes) // sample))[:sample]
    _MESH_JOB.update(pool=pool, shapes=shapes, verts=verts, tris=tris, nodes=nodes, xfs=xfs)
    with mp.get_context("fork").Pool(cores) as p:
        res = p.map(_mesh_cpu_worker, np.array_split(idx, cores), chunksize=1)
    wall = wall_span(res)
    cpu_s = sum(e - s for s, e in res)
    return dict({"value": len(idx) / wall, "unit": "contacts/s", "cores": cores, "kind": kind,
                 "sample": "%d of the %d contacts (strided), %d processes, %.1f CPU-s, wall %.2f s (tree built "
                           "beforehand in each process)" % (len(idx), len(shapes), cores, cpu_s, wall)}, **host)


def mesh_algorithmic_work(world, nodes, counts, stride=50):
    """FP32 ops of the mesh path (tree query, per-triangle b3CollideShapeAndShape,
    b3ClusterSolver::Run) from the op-counting oracle on every `stride`-th contact, scaled to
    the batch; and the algorithmic HBM bytes of one batch: per contact its shape index,
    shape record (48 B), transforms (96 B), triangle count and list words, three output
    manifolds (720 B) and manifold count; per (contact, triangle) work item the item (8 B,
    written and read), the triangle and its three vertices (64 B) and the per-triangle
    manifold (240 B, written by the narrowphase and read by the clustering); per hull contact
    the hull record once (12 V + 16 HE + 20 F + 24 B).  Tree nodes are not counted."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    pool, shapes, verts, tris, xfs = world
    oc = oracle_lib.oracle_counting()
    sa = np.arange(len(shapes), dtype=np.uint32)
    idx = sa[::stride]
    oc.lib.b3o_flop_reset()
    oracle_lib.collide_mesh(oc.lib, "b3o_", pool, shapes, verts, tris, nodes, idx, xfs[idx])
    flops = float(oc.lib.b3o_flop_count()) * len(sa) / len(idx)
    items = np.minimum(counts.astype(np.int64), 64)
    d = pool.desc
    hul = shapes["type"] == 2
    h = shapes["hull"][hul].astype(np.int64)
    hull_bytes = (12 * d[h, 1].astype(np.int64) + 16 * d[h, 3].astype(np.int64) + 20 * d[h, 5].astype(np.int64) + 24)
    n = len(shapes)
    byts = n * (4 + 48 + 96 + 4 + 720 + 4) + 4 * items.sum() + items.sum() * (16 + 64 + 480) + hull_bytes.sum()
    return {"flops": flops, "bytes": float(byts), "sample": len(idx)}


def mesh_run(world, steps, cpu):
    """Mixed shapes on a 256 x 256 heightfield (131,072 triangles): 100,000 fresh mesh
    contacts per batch (static-tree midphase, per-triangle narrowphase, clustering)."""
    import bounce_amd
    pool, shapes, verts, tris, xfs = world
    res = {"workload": "256x256 heightfield (131,072 triangles, value noise 2 m, cell 0.5 m); 100,000 shapes: "
                       "40% spheres r 0.1-0.5, 30% capsules L 0.2-1 r 0.1-0.3, 30% 8-32-point hulls 0.1-0.5 m; "
                       "uniform over the grid, resting gap N(0, 0.05 m); fresh contacts",
           "contacts": len(shapes)}
    if cpu is not None:
        res["cpu_baseline"] = cpu
    sa = np.arange(len(shapes), dtype=np.uint32)
    n = len(shapes)
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        mid = ctx.upload_mesh(verts, tris)
        # Host arrays (PCIe both ways, the b3amd_collide_mesh form).
        ctx.collide_mesh(mid, sa, xfs)
        t0 = time.perf_counter()
        for _ in range(steps):
            counts, lists, out, mcount = ctx.collide_mesh(mid, sa, xfs)
        host_wall = (time.perf_counter() - t0) / steps
        # Device-resident inputs and outputs (b3amd_collide_mesh_device).
        ds, dx = ctx.alloc(sa.nbytes), ctx.alloc(xfs.nbytes)
        do, dm = ctx.alloc(out.nbytes), ctx.alloc(mcount.nbytes)
        ds.upload(sa)
        dx.upload(xfs)
        ctx.collide_mesh_device(mid, n, ds, dx, do, dm)
        ctx.synchronize()
        ctx.reset_kernel_stats()
        ctx.set_profiling(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.collide_mesh_device(mid, n, ds, dx, do, dm)
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / steps
        ctx.set_profiling(False)
        kern = {k: ctx.kernel_stats(k) for k in ("mesh_midphase", "mesh_narrow", "mesh_cluster")}
        same = bool(np.array_equal(do.download(np.zeros_like(out)).view(np.uint8), out.view(np.uint8)))
        nodes = ctx.mesh_tree(mid)
    kms = {k: v[1] / max(1, v[0]) for k, v in kern.items()}
    dev_ms = sum(kms.values())
    work = mesh_algorithmic_work(world, nodes, counts)
    tflops = work["flops"] / (dev_ms / 1000.0) / 1e12
    gbs = work["bytes"] / (dev_ms / 1000.0) / 1e9
    res["roofline"] = {
        "bound": "valu", "achieved": tflops, "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": tflops / VALU_PEAK_TFLOPS, "unfused_frac": tflops / VALU_UNFUSED_TFLOPS,
        "hbm": {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS},
        "algorithmic_flops_per_batch": work["flops"], "algorithmic_bytes_per_batch": work["bytes"],
        "what": "the three mesh kernels together (their summed HIP-event times); FP32 ops counted by "
                "the op-counting oracle on every 50th contact, scaled; bytes as mesh_algorithmic_work "
                "states (tree nodes not counted)"}
    res.update({"ms_per_batch": wall * 1000.0, "contacts_per_s": n / wall,
                "host_arrays_ms_per_batch": host_wall * 1000.0, "host_arrays_contacts_per_s": n / host_wall,
                "kernels_ms": kms, "kernels_sum_ms": dev_ms, "device_equals_host_path": same,
                "triangles_per_contact": float(counts.mean()),
                "manifolds": int(mcount.sum()), "points": int(out["pointCount"].sum()),
                "touching_fraction": float((mcount > 0).mean()),
                "what": "b3amd_collide_mesh_device on resident inputs (value); b3amd_collide_mesh on host "
                        "arrays beside it"})
    if "cpu_baseline" in res:
        res["gpu_over_cpu"] = res["contacts_per_s"] / res["cpu_baseline"]["value"]
    return res


# ----------------------------------------------------------------- sphere/capsule pairs

CONVEX_PAIRS = 1 << 20
SERVING_WARMUP = 0.1  # seconds of arrivals before each measured Poisson run (see serving_run)
SERVING_RUNS = 3      # measured Poisson runs per rate; the run of median p99 is reported


def convex_run(world, steps, cpu):
    """The convex narrowphase (b3CollideSphereAndSphere ... b3CollideCapsuleAndHull,
    collide.cpp:172-223, GJK for the hull pairs) on 1,048,576 fresh sphere/capsule pairs of
    gen_convex_pairs, inputs and outputs resident: pairs/s of the batch and the convex
    kernel's time, with its roofline (FP32 ops from the op-counting oracle on every 64th
    pair; bytes: pair, transforms, two shape records, the manifold, and 12 B per vertex of a
    hull the GJK reads)."""
    import bounce_amd
    from bounce_amd.types import MANIFOLD_DTYPE
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    pool, shapes, pairs, xfs = world
    n = len(pairs)
    res = {"workload": "1,048,576 fresh pairs, equal shares of sphere-sphere, sphere-capsule, sphere-hull, "
                       "capsule-capsule, capsule-hull; spheres r 0.1-0.5, capsules r 0.1-0.3 L 0-1, hulls "
                       "8-32 points 0.2-1 m; placed along a random direction at touching distance + "
                       "N(-0.02, 0.08) m, 5 % nearly concentric", "pairs": n}
    if cpu is not None:
        res["cpu_baseline"] = cpu
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        dp, dx, do = ctx.alloc(pairs.nbytes), ctx.alloc(xfs.nbytes), ctx.alloc(n * MANIFOLD_DTYPE.itemsize)
        dp.upload(pairs)
        dx.upload(xfs)
        ctx.collide_device(n, dp, dx, do)
        ctx.synchronize()
        ctx.reset_kernel_stats()
        ctx.set_profiling(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.collide_device(n, dp, dx, do)
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / steps
        ctx.set_profiling(False)
        k = ctx.kernel_stats("convex")
        out = do.download(np.zeros(n, MANIFOLD_DTYPE))
    kms = k[1] / max(1, k[0])
    idx = np.arange(0, n, 64)
    oc = oracle_lib.oracle_counting()
    oc.lib.b3o_flop_reset()
    ref = oc.collide(pool, shapes, pairs[idx], xfs[idx])
    flops = float(oc.lib.b3o_flop_count()) * n / len(idx)
    same = bool(np.array_equal(out[idx].view(np.uint8), ref.view(np.uint8)))
    hv = np.where(shapes["type"] == 2, pool.desc[shapes["hull"], 1].astype(np.int64), 0)
    byts = float(n * (8 + 96 + 96 + 240) + 12 * hv.sum())
    tflops = flops / (kms / 1000.0) / 1e12
    gbs = byts / (kms / 1000.0) / 1e9
    res.update({"ms_per_batch": wall * 1000.0, "pairs_per_s": n / wall, "convex_kernel_ms": kms,
                "touching_fraction": float((out["pointCount"] > 0).mean()),
                "sample_equals_oracle": same,
                "roofline": {"bound": "valu", "achieved": tflops, "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": tflops / VALU_PEAK_TFLOPS, "unfused_frac": tflops / VALU_UNFUSED_TFLOPS,
                             "hbm": {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": gbs / HBM_PEAK_GBS},
                             "algorithmic_flops_per_batch": flops, "algorithmic_bytes_per_batch": byts},
                "what": "b3amd_collide_device on resident inputs; the convex kernel runs on its own stream "
                        "beside the (empty) hull kernels"})
    if "cpu_baseline" in res:
        res["gpu_over_cpu"] = res["pairs_per_s"] / res["cpu_baseline"]["value"]
    return res


# ----------------------------------------------------------------- large warm world

_WARM_JOB = {}
WARM_CPU_STEPS = 8
WARM_PROFILE_STEPS = 60


def _warm_cpu_worker(chunk):
    """The reference's chain over steps 0..WARM_CPU_STEPS-1 for the contacts in `chunk`;
    returns (cold-step seconds, warm-step seconds, last step's manifolds and caches)."""
    j = _WARM_JOB
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from bounce_amd.types import empty_caches
    from bounce_amd.workloads import warm_world_pair_inputs
    chk = oracle_lib.reference() if j["kind"] == "reference" else oracle_lib.oracle()
    w = j["world"]
    contacts = w["contacts"][chunk]
    caches = empty_caches(len(contacts))
    inputs = [warm_world_pair_inputs(w, xfs, sw, contacts) for xfs, sw in j["states"]]
    chk.collide(w["pool"], w["shapes"], inputs[0][0][:8], inputs[0][1][:8])  # load the library
    spans = []
    for pairs, pxf, psw in inputs:
        t0 = time.perf_counter()
        out = chk.collide(w["pool"], w["shapes"], pairs, pxf, sweeps=psw, caches=caches)
        spans.append((t0, time.perf_counter()))
    return spans[0], (spans[1][0], spans[-1][1]), out, caches


def warm_cpu_baseline(world, stride):
    """The reference (oracle/_ref; the oracle restatement where that is absent) stepping
    every `stride`-th contact of the warm world through WARM_CPU_STEPS steps with its
    caches carried, one process per core.  Runs before the GPU is initialised (forks)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from bounce_amd.workloads import warm_world_state
    kind = "reference" if oracle_lib.reference() is not None else "port"
    host = host_cpu()
    cores = host["processes"]
    idx = np.arange(0, len(world["contacts"]), stride)
    _WARM_JOB.update(kind=kind, world=world, states=[warm_world_state(world, s) for s in range(WARM_CPU_STEPS)])
    with mp.get_context("fork").Pool(cores) as p:
        res = p.map(_warm_cpu_worker, np.array_split(idx, cores), chunksize=1)
    cold = wall_span([r[0] for r in res])
    warm = wall_span([r[1] for r in res])
    cpu_s = sum(r[0][1] - r[0][0] + r[1][1] - r[1][0] for r in res)
    out = np.concatenate([r[2] for r in res])
    caches = np.concatenate([r[3] for r in res])
    return dict({"value": len(idx) * (WARM_CPU_STEPS - 1) / warm, "unit": "contacts/s (warm steps)",
                 "cores": cores, "kind": kind,
                 "cold_step_contacts_per_s": len(idx) / cold,
                 "sample": "every %d-th contact (%d of %d), steps 0-%d with caches carried, %d processes, "
                           "%.1f CPU-s, wall %.2f s cold + %.2f s warm" % (
                               stride, len(idx), len(world["contacts"]), WARM_CPU_STEPS - 1, cores, cpu_s,
                               cold, warm)}, **host), idx, out, caches


def warm_world_run(world, steps, cpu):
    """The large warm world: `steps` world steps (60 Hz) of b3amd_update_contacts on
    device-resident body state, each contact's feature cache kept on the device from step
    to step.  The per-step body states (the motion repeats every `period` steps, so
    period + 1 distinct states) are uploaded before the timed run."""
    import bounce_amd
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from util import defined_hull_cache, HULL_CACHE_FIELDS
    from bounce_amd.types import CACHE_DTYPE, MANIFOLD_DTYPE, empty_caches
    from bounce_amd.workloads import warm_world_pair_inputs, warm_world_state, warm_world_state_index
    w = world
    B, n = len(w["body_shapes"]), len(w["contacts"])
    nstates = w["period"] + 1
    res = {"workload": "%d hulls of 8-64 points (semi-axes U[0.25,0.5] m, uniform orientations) in %d x %d "
                       "stacks of up to %d in a 50 x 50 m container (floor and 4 walls); contact list = fat-AABB "
                       "overlaps (0.2 m); synthetic motion: oscillation of 2 mm / 0.005 rad (10%%: 2 cm / "
                       "0.05 rad) at 0.5-2 Hz, period %d steps" % (B - w["statics"], w["columns"], w["columns"],
                                                                    w["layers"], w["period"]),
           "bodies": B, "contacts": n, "steps": steps}
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(w["pool"])
        ctx.upload_shapes(w["shapes"])
        d_shapes, d_contacts = ctx.alloc(4 * B), ctx.alloc(8 * n)
        d_caches, d_out = ctx.alloc(CACHE_DTYPE.itemsize * n), ctx.alloc(MANIFOLD_DTYPE.itemsize * n)
        d_shapes.upload(w["body_shapes"])
        d_contacts.upload(w["contacts"])
        d_x, d_s = ctx.alloc(48 * B * nstates), ctx.alloc(32 * B * nstates)
        for i in range(nstates):
            xfs, sw = warm_world_state(w, i)
            d_x.view(48 * B * i, 48 * B).upload(xfs)
            d_s.view(32 * B * i, 32 * B).upload(sw)

        def step(s):
            i = warm_world_state_index(w, s)
            ctx.update_contacts(B, d_shapes, d_x.view(48 * B * i, 48 * B), n, d_contacts, d_out,
                                d_body_sweeps=d_s.view(32 * B * i, 32 * B), d_caches=d_caches)

        errs = []

        def fp64_error(s):
            out = d_out.download(np.zeros(n, MANIFOLD_DTYPE))
            xfs, sw = warm_world_state(w, s)
            errs.append((out, warm_world_pair_inputs(w, xfs, sw)[1]))
            return out

        # Parity pass: the first WARM_CPU_STEPS steps, compared with the reference's chain; the
        # statistics counters (b3_convexCalls, b3_convexCacheHits) count its warm steps.
        d_caches.upload(empty_caches(n))
        step(0)
        ctx.synchronize()
        fp64_error(0)
        ctx.set_counting(True)
        ctx.reset_counters()
        for s in range(1, WARM_CPU_STEPS):
            step(s)
        ctx.synchronize()
        warm = ctx.counters()
        ctx.set_counting(False)
        if cpu is not None:
            stats, idx, want_out, want_caches = cpu
            out = d_out.download(np.zeros(n, MANIFOLD_DTYPE))[idx]
            caches = d_caches.download(np.zeros(n, CACHE_DTYPE))[idx]
            man_eq = (out.view(np.uint8).reshape(len(idx), -1) == want_out.view(np.uint8).reshape(len(idx), -1)).all(1)
            dc, dw = defined_hull_cache(caches), defined_hull_cache(want_caches)
            cache_eq = np.all([dc[f] == dw[f] for f in HULL_CACHE_FIELDS], axis=0)
            res["sampled_parity"] = {"contacts": int(len(idx)), "step": WARM_CPU_STEPS - 1,
                                     "manifolds_equal": int(man_eq.sum()), "caches_equal": int(cache_eq.sum()),
                                     "against": stats["kind"]}
            res["cpu_baseline"] = stats
        # Timed run from empty caches: the cold step, then the warm steps.
        d_caches.upload(empty_caches(n))
        ctx.synchronize()
        t0 = time.perf_counter()
        step(0)
        ctx.synchronize()
        t1 = time.perf_counter()
        t2 = time.perf_counter()
        for s in range(1, steps):
            step(s)
        ctx.synchronize()
        t3 = time.perf_counter()
        last = fp64_error(steps - 1)
        # Per-kernel times from further warm steps with events around each launch (kept out
        # of the timed steps: the events add host work and gaps between launches).
        ctx.reset_kernel_stats()
        ctx.set_profiling(True)
        for s in range(steps, steps + WARM_PROFILE_STEPS):
            step(s)
        ctx.synchronize()
        ctx.set_profiling(False)
        kern = {k: ctx.kernel_stats(k) for k in ("gather", "cache_hull", "sat_hull", "contact_hull")}
    warm_s = t3 - t2
    calls, hits = warm["convex_calls"], warm["convex_cache_hits"]
    res.update({
        "cold_step_ms": (t1 - t0) * 1000.0,
        "warm_ms_per_step": warm_s * 1000.0 / (steps - 1),
        "warm_contacts_per_s": n * (steps - 1) / warm_s,
        "warm_steps_per_s": (steps - 1) / warm_s,
        "cache_hit_fraction": hits / max(1, calls),
        "cache_hit_fraction_steps": [1, WARM_CPU_STEPS - 1],
        "kernels_ms": {k: v[1] / max(1, v[0]) for k, v in kern.items()},
        "touching_fraction_last_step": float((last["pointCount"] > 0).mean()),
        "points_last_step": int(last["pointCount"].sum()),
        "what": "b3amd_update_contacts per step on device-resident body transforms and sweeps; caches and "
                "manifolds stay in HBM; statistics counters and kernel events off in the timed steps "
                "(kernels_ms from %d further steps with events)" % WARM_PROFILE_STEPS})
    outs = np.concatenate([o for o, _ in errs])
    xfs = np.concatenate([x for _, x in errs])
    res["separation_error_fp64"] = separation_error_fp64(outs, xfs)
    res["separation_error_fp64"]["steps"] = [0, steps - 1]
    if "cpu_baseline" in res:
        res["gpu_over_cpu"] = res["warm_contacts_per_s"] / res["cpu_baseline"]["value"]
    return res


# ----------------------------------------------------------------- C5: the rain-in world

RAIN_PATCH = 6  # columns per side of each CPU-baseline patch world


def _rain_patch_bodies(world, i0, j0, k):
    """Bodies of the k x k columns from (i0, j0), every layer, plus the statics."""
    C, ns = world["columns"], world["statics"]
    n = len(world["body_shapes"]) - ns
    cells = np.arange(n)
    i, j = cells % C, (cells // C) % C
    sel = cells[(i >= i0) & (i < i0 + k) & (j >= j0) & (j < j0 + k)] + ns
    return np.concatenate([np.arange(ns), sel])


def _rain_cpu_worker(patch):
    """The reference's own b3World (contact manager, broadphase, narrowphase) stepping one
    patch of the rain-in for every step; returns (start, end, narrowphase calls)."""
    j = _RAIN_JOB
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from bounce_amd.workloads import rain_state
    w = j["world"]
    chk = oracle_lib.reference() if j["kind"] == "reference" else oracle_lib.oracle()
    prefix = "b3ref_kworld_" if j["kind"] == "reference" else "b3o_kworld_"
    bodies = _rain_patch_bodies(w, patch[0], patch[1], RAIN_PATCH)
    spawn = w["spawn"][bodies]
    kw = oracle_lib.KWorld(chk.lib, prefix, w["pool"], w["shapes"])
    # inputs first (untimed): the states of the patch's bodies per step
    steps = j["steps"]
    states = [rain_state(w, s, bodies) for s in range(steps + 1)]
    order = []
    t0 = time.perf_counter()
    calls = 0
    st = np.nonzero(spawn < 0)[0]
    kw.add_bodies(w["body_shapes"][bodies[st]], np.zeros(len(st), np.uint32), states[0][0][st], states[0][1][st])
    order.extend(st.tolist())
    kw.find_new_contacts()
    for s in range(steps):
        new = np.nonzero(spawn == s)[0]
        if len(new):
            kw.add_bodies(w["body_shapes"][bodies[new]], np.full(len(new), 2, np.uint32), states[s][0][new],
                          states[s][1][new])
            order.extend(new.tolist())
            kw.find_new_contacts()
        calls += int(kw.update_contacts()[1])
        present = np.array(order)
        mv = np.nonzero(spawn[present] >= 0)[0]
        kw.move_bodies(mv, states[s + 1][0][present[mv]], states[s + 1][1][present[mv]])
        kw.find_new_contacts()
    t1 = time.perf_counter()
    kw.close()
    return t0, t1, calls


def rain_cpu_baseline(world, steps):
    """The reference's own b3World over the rain-in, restricted to disjoint patches of
    RAIN_PATCH x RAIN_PATCH columns (every layer, with the container), one patch world per
    process, all steps: UpdateContacts with its narrowphase, SynchronizeShapes and
    FindNewContacts, as b3World::Step runs them.  Runs before the GPU is initialised."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    kind = "reference" if oracle_lib.reference() is not None else "port"
    host = host_cpu()
    cores = host["processes"]
    C = world["columns"]
    patches = [(i, j) for j in range(0, C - RAIN_PATCH + 1, RAIN_PATCH) for i in range(0, C - RAIN_PATCH + 1, RAIN_PATCH)]
    patches = patches[:: max(1, len(patches) // cores)][:cores]
    _RAIN_JOB.update(world=world, steps=steps, kind=kind)
    with mp.get_context("fork").Pool(cores) as p:
        res = p.map(_rain_cpu_worker, patches, chunksize=1)
    wall = wall_span([(r[0], r[1]) for r in res])
    calls = sum(r[2] for r in res)
    bodies = sum(len(_rain_patch_bodies(world, i, j, RAIN_PATCH)) for i, j in patches)
    return dict({"value": calls / wall, "unit": "narrowphase contacts/s (whole contact-manager step)",
                 "cores": cores, "kind": kind,
                 "sample": "%d patch worlds of %d x %d columns (%d bodies in all), %d steps each, one process per "
                           "patch: the reference's b3World contact manager, broadphase and narrowphase; %d "
                           "narrowphase calls, wall %.1f s" % (len(patches), RAIN_PATCH, RAIN_PATCH, bodies, steps,
                                                               calls, wall)}, **host)


RAIN_PROFILE_EVERY = 10


def rain_world_run(lib, world, steps, cpu, sample=2000):
    """The rain-in (configs[4] as specified): hulls spawn (b3Body::CreateShape) and fall into
    the container; each step runs, on the device, what b3World::Step runs around the solver:
    FindNewContacts for the new shapes, UpdateContacts (contacts destroyed when their fat
    AABBs part, the narrowphase with each contact's cache for the others), the move with
    SynchronizeShapes, and FindNewContacts (broadphase pairs, AddPair).  The kinematic motion
    stands in for the solver: its states are computed on the host and uploaded outside the
    timed regions."""
    import bounce_amd
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from util import defined_hull_cache, HULL_CACHE_FIELDS
    from bounce_amd.types import SWEEP_DTYPE, MANIFOLD_DTYPE
    from bounce_amd.workloads import rain_state
    w = world
    B, ns = len(w["body_shapes"]), w["statics"]
    spawn = w["spawn"]
    fall_steps = int(np.ceil(w["t_fall"] / w["dt"])) + 2
    res = {"workload": "%d hulls of 8-64 points raining into a 50 x 50 m container (floor and 4 walls): each "
                       "spawns %.0f m above its resting place in its column (layer k at step %d k + a column "
                       "offset), falls under %.1f m/s^2 turning to its resting orientation, and rests; %d steps "
                       "at 60 Hz; kinematic motion (no solver)" % (B - ns, w["fall_height"], w["interval"],
                                                                   w["gravity"], steps),
           "bodies": B, "steps": steps}
    # Host-side body state, advanced per step for the bodies in motion only.
    xf, sw, q = rain_state(w, 0)
    changing = lambda s: np.nonzero((spawn >= 0) & (spawn <= s + 1) & (s <= spawn + fall_steps))[0]
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(w["pool"])
        ctx.upload_shapes(w["shapes"])
        d_xf, d_q, d_sw, d_ids = ctx.alloc(48 * B), ctx.alloc(32 * B), ctx.alloc(72 * B), ctx.alloc(4 * B)
        with ctx.broadphase() as bp, ctx.contact_manager(bp) as m:
            flags = np.where(np.arange(B) < ns, 0 | 4, 2 | 4).astype(np.uint32)
            m.set_bodies(flags)
            shape_body, shape_proxy = [], []
            st = np.nonzero(spawn < 0)[0]
            d_xf.upload(xf)
            shape_proxy.extend(m.add_shapes(w["body_shapes"][st], st, d_xf).tolist())
            shape_body.extend(st.tolist())
            m.find_new_contacts()
            ctx.reset_kernel_stats()
            step_ms, created, destroyed, collided, listed, profiled = [], [], [], [], [], []
            ctx.set_counting(True)
            ctx.reset_counters()
            for s in range(steps):
                # Every RAIN_PROFILE_EVERY-th step brackets its launches with events for the
                # per-kernel times; the step times and rates come from the other steps.
                prof = s % RAIN_PROFILE_EVERY == 0
                ctx.set_profiling(prof)
                profiled.append(prof)
                d_xf.upload(xf)
                d_q.upload(q)
                ctx.synchronize()
                t0 = time.perf_counter()
                new = np.nonzero(spawn == s)[0]
                c = 0
                if len(new):
                    shape_proxy.extend(m.add_shapes(w["body_shapes"][new], new, d_xf).tolist())
                    shape_body.extend(new.tolist())
                    c += m.find_new_contacts()
                stats = m.update_contacts(d_xf, d_q)
                ctx.synchronize()
                t1 = time.perf_counter()
                listed.append(m.count())
                ch = changing(s)
                x1, s1, q1 = rain_state(w, s + 1, ch)
                xf[ch], sw[ch], q[ch] = x1, s1, q1
                d_xf.upload(xf)
                d_sw.upload(sw)
                ids = np.nonzero(spawn[np.array(shape_body)] >= 0)[0].astype(np.uint32)
                d_ids.upload(ids)
                ctx.synchronize()
                t2 = time.perf_counter()
                m.synchronize(len(ids), d_ids, d_xf, d_sw)
                c += m.find_new_contacts()
                ctx.synchronize()
                t3 = time.perf_counter()
                step_ms.append(((t1 - t0) + (t3 - t2)) * 1000.0)
                created.append(c)
                destroyed.append(int(stats[0]))
                collided.append(int(stats[1]))
            ctx.set_profiling(False)
            counters = ctx.counters()
            ctx.set_counting(False)
            kern = {}
            for k in ("bp_move", "bp_lists", "bp_query", "cm_sync", "cm_filter", "cm_create", "cm_classify",
                      "cm_scatter", "cm_overlap", "gather", "cache_hull", "sat_hull", "contact_hull", "convex"):
                n_, t_ = ctx.kernel_stats(k)
                if n_:
                    kern[k] = t_ / max(1, int(np.sum(profiled)))
            pairs, born, overlap, out, caches = m.read()
            fat = bp.read(0, bp.info()[1])[0][np.array(shape_proxy)]
    timed_steps = ~np.array(profiled)
    step_ms = np.array(step_ms)[timed_steps]
    total_s = step_ms.sum() / 1000.0
    res.update({
        "ms_per_step": float(step_ms.mean()),
        "ms_per_step_p50_p99_max": [float(np.percentile(step_ms, 50)), float(np.percentile(step_ms, 99)),
                                    float(step_ms.max())],
        "contacts_per_s": float(np.sum(np.array(collided)[timed_steps]) / total_s),
        "steps_per_s": len(step_ms) / total_s,
        "timed_steps": int(len(step_ms)),
        "contacts_created": int(np.sum(created)), "contacts_destroyed": int(np.sum(destroyed)),
        "narrowphase_calls": int(np.sum(collided)),
        "contacts_at_end": int(listed[-1]), "touching_at_end": int((out["pointCount"] > 0).sum()),
        "cache_hit_fraction": counters["convex_cache_hits"] / max(1, counters["convex_calls"]),
        "kernel_ms_per_step": kern,
        "what": "per step: b3amd_contact_manager_add_shapes + find_new_contacts for the new shapes, "
                "update_contacts (destroy, narrowphase with each contact's cache), synchronize (SynchronizeShapes "
                "+ MoveProxy of every shape present) and find_new_contacts (FindPairs + AddPair), all on the "
                "device, host wall clock including their synchronisations; body states uploaded outside it; "
                "rates over the steps run without kernel events, kernel_ms_per_step from every %dth step, "
                "which records them" % RAIN_PROFILE_EVERY,
    })
    res["timeline"] = {"every": 50, "contacts": [int(x) for x in listed[::50]],
                       "created": [int(x) for x in np.add.reduceat(created, np.arange(0, steps, 50))],
                       "destroyed": [int(x) for x in np.add.reduceat(destroyed, np.arange(0, steps, 50))]}
    # Sampled parity: chains of the reference's narrowphase from each sampled contact's
    # creation to the last step, with its cache carried, against the device's.
    if cpu is not None:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        ref = oracle_lib.reference() if cpu["kind"] == "reference" else oracle_lib.oracle()
        from bounce_amd.types import empty_caches
        idx = np.arange(0, len(pairs), max(1, len(pairs) // sample))
        sb = np.array(shape_body)
        bodies = sb[pairs[idx]]
        need = np.unique(bodies)
        where = {b: i for i, b in enumerate(need)}
        loc = np.vectorize(where.get)(bodies)
        cch = empty_caches(len(idx))
        got = np.zeros(len(idx), MANIFOLD_DTYPE)
        b0 = born[idx]
        for s_ in range(int(b0.min()), steps):
            act = np.nonzero(b0 <= s_)[0]
            xs, _, qs = rain_state(w, s_, need)
            pr = w["body_shapes"][bodies[act]].astype(np.uint32)
            pxf = np.ascontiguousarray(xs[loc[act]].reshape(-1, 2, 12))
            psw = np.ascontiguousarray(qs[loc[act]].reshape(-1, 4, 4))
            sub = np.ascontiguousarray(cch[act])
            got[act] = ref.collide(w["pool"], w["shapes"], pr, pxf, sweeps=psw, caches=sub)
            cch[act] = sub
        man_eq = (got.view(np.uint8).reshape(len(idx), -1) == out[idx].view(np.uint8).reshape(len(idx), -1)).all(1)
        dc, dw = defined_hull_cache(caches[idx]), defined_hull_cache(cch)
        cache_eq = np.all([dc[f] == dw[f] for f in HULL_CACHE_FIELDS], axis=0)
        res["sampled_parity"] = {"contacts": int(len(idx)), "step": steps - 1, "manifolds_equal": int(man_eq.sum()),
                                 "caches_equal": int(cache_eq.sum()), "against": cpu["kind"],
                                 "what": "each sampled contact's narrowphase chain replayed by the reference from "
                                         "its creation step with its cache carried"}
        res["cpu_baseline"] = cpu
        res["gpu_over_cpu"] = res["contacts_per_s"] / cpu["value"]
    # Contact set check: after the last UpdateContacts the list holds exactly the shape pairs
    # whose fat AABBs overlap (not both static), from the broadphase's own fat AABBs.
    res["contact_set_equals_fat_aabb_overlaps"] = rain_contact_set_check(fat, pairs, np.array(shape_body) < ns)
    res["separation_error_fp64"] = rain_fp64_error(w, out, pairs, shape_body, steps - 1)
    return res


def rain_contact_set_check(fat, pairs, static):
    """Whether the contact list's shape pairs are exactly the pairs of shapes whose fat AABBs
    overlap (b3TestOverlap on float32 boxes), excluding static-static pairs; brute force
    through a k-d tree over the box centres (the reach covers the largest box)."""
    from scipy.spatial import cKDTree
    lo, hi = fat[:, :3], fat[:, 3:]
    dyn = np.nonzero(~static)[0]
    c = 0.5 * (lo[dyn].astype(np.float64) + hi[dyn])
    half = 0.5 * (hi[dyn].astype(np.float64) - lo[dyn]).max()
    pr = cKDTree(c).query_pairs(2.0 * half + 1e-3, p=np.inf, output_type="ndarray")
    pr = dyn[pr]
    ok = np.all((lo[pr[:, 0]] <= hi[pr[:, 1]]) & (hi[pr[:, 0]] >= lo[pr[:, 1]]), axis=1)
    want = [pr[ok]]
    for a in np.nonzero(static)[0]:
        m = np.all((lo[dyn] <= hi[a]) & (hi[dyn] >= lo[a]), axis=1)
        want.append(np.stack([np.full(int(m.sum()), a), dyn[m]], 1))
    want = np.sort(np.concatenate(want), axis=1)
    got = np.sort(pairs.astype(np.int64), axis=1)
    key = lambda p: np.unique(p[:, 0] * (1 << 32) + p[:, 1])
    return bool(len(got) == len(key(got)) and np.array_equal(key(got), key(want)))


def rain_fp64_error(world, out, pairs, shape_body, s):
    from bounce_amd.workloads import rain_state
    sb = np.array(shape_body)
    bodies = sb[pairs]
    xs, _, _ = rain_state(world, s)
    return separation_error_fp64(out, np.ascontiguousarray(xs[bodies]))


# ----------------------------------------------------------------- C1: the box pyramid

def pyramid_run(lib):
    """The box pyramid (210 unit boxes on a static ground) replayed step by step from the
    trace of the reference's own world: one b3amd_collide call per step on host arrays
    (the plain drop-in form), feature caches carried by contact id, each step's output
    digest compared with the reference's."""
    import bounce_amd
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from util import (PYRAMID_STEPS, CacheChain, load_golden, pyramid_inputs, pyramid_shapes, step_digest)
    t = load_golden("pyramid_trace.npz")
    chain = CacheChain()
    equal = contacts = points = 0
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(t["pool"])
        ctx.upload_shapes(pyramid_shapes(t["body_xfs"].shape[1]))
        inputs = [pyramid_inputs(t, s) for s in range(PYRAMID_STEPS)]
        ctx.collide(inputs[-1][1], inputs[-1][2])  # warm-up (scratch allocation)
        wall = 0.0
        for s, (ids, pairs, xfs, sweeps) in enumerate(inputs):
            caches = chain.gather(ids)
            t0 = time.perf_counter()
            out = ctx.collide(pairs, xfs, sweeps=sweeps, caches=caches)
            wall += time.perf_counter() - t0
            chain.scatter(ids, caches)
            equal += int(np.array_equal(step_digest(out, caches), t["digests"][s]))
            contacts += len(ids)
            points += int(out["pointCount"].sum())
    res = {"steps": PYRAMID_STEPS, "steps_equal_to_reference": equal, "contacts": contacts, "points": points,
           "ms_per_step": wall * 1000.0 / PYRAMID_STEPS,
           "what": "b3amd_collide per step on host arrays (H2D, kernels, D2H, synchronise); inputs from "
                   "the reference world's trace (tests/golden/pyramid_trace.npz)"}
    res["whole_step"] = pyramid_world_run(t)
    return res


def pyramid_world_run(t):
    """The same pyramid stepped end to end by the library alone (b3amd_host_world: islands and
    solver on the host, broadphase, contact manager and narrowphase on the device), each step
    compared with the reference world's trace; beside it the reference's own b3World::Step
    timed on one host core where it is built."""
    import bounce_amd
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from bounce_amd.types import SHAPE_DTYPE, HULL
    from util import PYRAMID_STEPS
    levels = 20
    f32 = np.float32
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(t["pool"])
        shapes = np.zeros(2, SHAPE_DTYPE)
        shapes["type"] = HULL
        shapes["hull"] = [0, 1]
        ctx.upload_shapes(shapes)
        with ctx.host_world() as w:
            w.create_shape(w.create_body(type=0), 0)
            for l in range(levels):
                for j in range(levels - l):
                    x = f32(2.1) * f32(j) + f32(l) * f32(1.0) - f32(0.5) * f32(levels) * f32(2.0)
                    w.create_shape(w.create_body(type=2, position=(x, f32(3.0) * f32(l + 1), 0.0)), 1, density=1.0)
            equal, ms = 0, []
            for s in range(PYRAMID_STEPS - 1):
                t0 = time.perf_counter()
                w.step(1.0 / 60.0, 8, 2)
                ms.append((time.perf_counter() - t0) * 1000.0)
                st = w.bodies()
                equal += int(np.array_equal(st["xf"].view(np.uint32), t["body_xfs"][s + 1].view(np.uint32)))
            contacts = w.contact_count()
    res = {"steps": PYRAMID_STEPS - 1, "steps_equal_to_reference": equal, "ms_per_step": float(np.mean(ms)),
           "ms_per_step_p50_p99": [float(np.percentile(ms, 50)), float(np.percentile(ms, 99))],
           "contacts_at_end": contacts,
           "what": "b3amd_host_world_step: b3World::Step with the islands and sequential-impulse solver on the "
                   "host and the broadphase, contact manager and narrowphase on the device; every body transform "
                   "compared with the reference world's after each step"}
    ref = oracle_lib.reference()
    if ref is not None:
        world = oracle_lib.RefPyramid(ref.lib)
        t0 = time.perf_counter()
        for _ in range(PYRAMID_STEPS - 1):
            world.step()
        res["reference_ms_per_step"] = (time.perf_counter() - t0) * 1000.0 / (PYRAMID_STEPS - 1)
        res["reference_what"] = "the reference's own b3World::Step on one host core (its driver also records each " \
                                "step's narrowphase inputs and outputs)"
        world.close()
    return res


# ----------------------------------------------------------------- main

TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "sat_hull_1M_pairs_hbm_traffic.md")


def measured_traffic(n_pairs):
    """HBM bytes per SAT launch from the committed FETCH_SIZE/WRITE_SIZE passes (measured on
    the 1,048,576-pair workload only; None for other sizes)."""
    if n_pairs != 1 << 20 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch sat_hull:"):
            return float(line.split(":")[1])
    return None


def serving_run(lib, worlds_n, rate, duration, seed=4096):
    """Serving workload: `worlds_n` worlds (10-2000 bodies, log-uniform; boxes, spheres and
    8-32-point hulls in stacks and piles, contacts from fat-AABB overlaps), each step one
    request.  Measures the flood capacity (two warm steps of every world submitted at once)
    with one lane and with three overlapping lanes, the latency of Poisson arrivals at
    `rate` steps/s and at half the capacity for `duration` seconds each, and beside them the
    plain drop-in: one b3amd_collide call on host arrays per world step.  The measured
    Poisson runs follow one unmeasured run of 3 x SERVING_WARMUP seconds at `rate` (the first
    run after the floods pays one-time costs at its start: its request records' host memory
    is touched for the first time), and each opens with SERVING_WARMUP seconds of arrivals at
    the same rate whose latencies are not counted; the warm-up run's and the prefixes' p99 and
    max are reported beside the steady-state figures.  Each rate runs SERVING_RUNS times: a
    stall of several milliseconds now and then (the box's GPU or host, not any one batch's
    work) lifts one run's p99 above the others, so the run of median p99 is the figure and all
    runs' p50/p99/max are listed."""
    import bounce_amd
    from bounce_amd.types import empty_caches
    from bounce_amd.workloads import gen_serving_worlds, world_pair_inputs
    t = time.perf_counter()
    pool, shapes, worlds = gen_serving_worlds(lib, seed, worlds_n, steps=4)
    total_pairs = sum(len(w["pairs"]) for w in worlds)
    total_bodies = sum(len(w["body_shapes"]) for w in worlds)
    log("serving workload: %d worlds, %d bodies, %d contacts in %.1f s"
        % (worlds_n, total_bodies, total_pairs, time.perf_counter() - t))
    rng = np.random.default_rng(seed)
    out = {"worlds": worlds_n, "bodies": total_bodies, "contacts": total_pairs,
           "workload": "worlds of 10-2000 bodies (log-uniform): boxes, spheres, 8-32-point hulls in stacks "
                       "(1-6 high) and piles (2x2 or 3x3 columns of 1-3) on a static ground; contacts = fat-AABB "
                       "overlaps (0.2 m), %.2f per body; per-step position jitter 1 mm; caches and manifolds "
                       "resident, body transforms sent per step" % (total_pairs / total_bodies)}
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(shapes)
        with bounce_amd.Server(ctx, 1 << 16, 1 << 16, total_pairs + 4096) as srv:
            ids = np.array([srv.admit_world(w["body_shapes"], w["pairs"], 1000.0) for w in worlds], np.uint32)
            # Cold step (empty caches) of every world, then the flood capacity on warm steps.
            srv.replay(ids, np.zeros(worlds_n), [w["xfs"][0] for w in worlds])
            order = np.concatenate([ids, ids])
            steps = np.concatenate([np.ones(worlds_n, int), np.full(worlds_n, 2)])
            flood = [worlds[i]["xfs"][s] for i, s in zip(order, steps)]
            for lanes in (1, 3):
                srv.configure(batch_pairs=1 << 16, max_inflight=lanes)
                lat = srv.replay(order, np.zeros(len(order)), flood)
                capacity = len(order) / (lat.max() / 1000.0)
                out["capacity_steps_per_s_%d_lane%s" % (lanes, "s" if lanes > 1 else "")] = capacity
                log("serving capacity, %d lane(s): %.0f steps/s" % (lanes, capacity))
            out["capacity_steps_per_s"] = capacity
            out["capacity_contacts_per_s"] = capacity * total_pairs / worlds_n
            # One unmeasured Poisson run first: the first run after the floods pays one-time
            # costs of its request records (host memory first touched), milliseconds that land
            # at its start.
            nwu = int(rate * 3 * SERVING_WARMUP)
            twu = np.cumsum(rng.exponential(1.0 / rate, nwu))
            wwu, swu = rng.integers(0, worlds_n, nwu), rng.integers(1, 4, nwu)
            lat_wu = srv.replay(ids[wwu], twu, [worlds[w]["xfs"][s] for w, s in zip(wwu, swu)])
            out["warmup_run"] = {"rate_steps_per_s": rate, "requests": nwu,
                                 "p99_ms": float(np.percentile(lat_wu, 99)), "max_ms": float(lat_wu.max())}
            for name, r in (("at_rate", rate), ("at_half_capacity", 0.5 * capacity)):
                runs = []
                for _ in range(SERVING_RUNS):
                    nw = int(r * SERVING_WARMUP)
                    n = int(r * duration)
                    times = np.cumsum(rng.exponential(1.0 / r, nw + n))
                    wsel = rng.integers(0, worlds_n, nw + n)
                    ssel = rng.integers(1, 4, nw + n)
                    before = srv.stats()
                    lat_all = srv.replay(ids[wsel], times, [worlds[w]["xfs"][s] for w, s in zip(wsel, ssel)])
                    after = srv.stats()
                    batches = after["batches"] - before["batches"]
                    lat, t_run = lat_all[nw:], times[nw:] - times[nw]
                    runs.append({
                        "requests": n,
                        "p50_ms": float(np.percentile(lat, 50)),
                        "p99_ms": float(np.percentile(lat, 99)),
                        "max_ms": float(lat.max()),
                        "prefix": {"seconds": SERVING_WARMUP, "requests": nw,
                                   "p99_ms": float(np.percentile(lat_all[:nw], 99)) if nw else None,
                                   "max_ms": float(lat_all[:nw].max()) if nw else None},
                        "achieved_steps_per_s": n / float((t_run + lat / 1000.0).max()),
                        "batches": batches,
                        "mean_batch_contacts": (after["pairs"] - before["pairs"]) / max(1, batches),
                        "device_ms_per_batch": (after["gpu_ms"] - before["gpu_ms"]) / max(1, batches),
                        "pack_ms_per_batch": (after["pack_ms"] - before["pack_ms"]) / max(1, batches),
                    })
                # The run of median p99 stands for the rate; every run is listed.
                med = sorted(runs, key=lambda x: x["p99_ms"])[len(runs) // 2]
                out[name] = dict(med, rate_steps_per_s=r, runs=[{k: x[k] for k in ("p50_ms", "p99_ms", "max_ms")}
                                                                 for x in runs])
                log("serving %s: %.0f steps/s, p50 %.3f ms, p99 %.3f ms (runs' p99: %s), %d batches"
                    % (name, r, out[name]["p50_ms"], out[name]["p99_ms"],
                       ", ".join("%.3f" % x["p99_ms"] for x in runs), med["batches"]))
        # Plain drop-in: one synchronous b3amd_collide per world step on host arrays, caches
        # carried per world, for 256 sampled worlds (warm steps 1-3).
        pick = rng.choice(worlds_n, 256, replace=False)
        caches = {int(k): empty_caches(len(worlds[k]["pairs"])) for k in pick}
        for k in pick:
            p, x = world_pair_inputs(worlds[k], 0)
            ctx.collide(p, x, caches=caches[int(k)])
        plain = []
        for step in (1, 2, 3):
            for k in pick:
                p, x = world_pair_inputs(worlds[k], step)
                t0 = time.perf_counter()
                ctx.collide(p, x, caches=caches[int(k)])
                plain.append(time.perf_counter() - t0)
        plain = np.array(plain) * 1000.0
        out["plain_drop_in"] = {"p50_ms": float(np.percentile(plain, 50)), "p99_ms": float(np.percentile(plain, 99)),
                                "steps_per_s": 1000.0 / float(plain.mean()),
                                "what": "one b3amd_collide per world step (host arrays in and out, synchronous), "
                                        "256 sampled worlds x 3 warm steps, one caller thread"}
        log("serving plain drop-in: %.0f steps/s, p50 %.3f ms, p99 %.3f ms"
            % (out["plain_drop_in"]["steps_per_s"], out["plain_drop_in"]["p50_ms"], out["plain_drop_in"]["p99_ms"]))
    out["latency"] = "narrowphase only: request submitted (scheduled arrival) -> manifolds in HBM"
    return out


def other_forms(ctx, n, dp, dx, do, steps, eager_out):
    """The same batch in the two other forms a caller runs it: with feature caches that start
    empty (b3ConvexContact's new-contact state, convex_contact.cpp:35-36; the batch reads and
    writes 36 B of cache per pair; the caches are re-emptied by a fill kernel before every
    batch, inside the timed region) and as a captured HIP graph replayed `steps` times.  Both
    must give the eager batch's manifolds."""
    from bounce_amd.types import MANIFOLD_DTYPE
    res = {}
    caches = ctx.cache_alloc(n)
    try:
        caches.reset()
        ctx.collide_device(n, dp, dx, do, d_caches=caches)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            caches.reset()
            ctx.collide_device(n, dp, dx, do, d_caches=caches)
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        same = do.download(np.zeros(n, MANIFOLD_DTYPE)).tobytes() == eager_out.tobytes()
        res["empty_caches"] = {"value": n / (ms / 1000.0), "unit": "pairs/s", "ms_per_step": ms,
                               "manifolds_equal_eager": same}
    finally:
        caches.free()
    with ctx.capture(n, dp, dx, do) as g:
        g.launch()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.launch()
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
    same = do.download(np.zeros(n, MANIFOLD_DTYPE)).tobytes() == eager_out.tobytes()
    res["graph_replay"] = {"value": n / (ms / 1000.0), "unit": "pairs/s", "ms_per_step": ms,
                           "manifolds_equal_eager": same}
    log("empty-cache form %.3f ms per batch, graph replay %.3f ms per batch"
        % (res["empty_caches"]["ms_per_step"], ms))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=1 << 20)
    ap.add_argument("--pool", type=int, default=65536)
    ap.add_argument("--cpu-sample", type=int, default=262144)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-pyramid", action="store_true")
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--no-convex", action="store_true", help="skip the sphere/capsule pair batch")
    ap.add_argument("--no-plain", action="store_true", help="skip the host-array (PCIe) drop-in timing")
    ap.add_argument("--no-warm", action="store_true", help="skip the large warm world")
    ap.add_argument("--warm-steps", type=int, default=600)
    ap.add_argument("--no-rain", action="store_true", help="skip the rain-in world")
    ap.add_argument("--rain-steps", type=int, default=600)
    ap.add_argument("--serving-worlds", type=int, default=4096)
    ap.add_argument("--serving-rate", type=float, default=240000.0)
    ap.add_argument("--serving-seconds", type=float, default=0.5)
    ap.add_argument("--tuning", default="auto", metavar="auto|W,K",
                    help="SAT launch shape: autotune on the batch, or waves-per-SIMD,tail-chunk")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run one device-to-device copy of BYTES (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the narrowphase has no cross-GPU exchange per step "
              "(pairs shard by contiguous ranges); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import bounce_amd
    from bounce_amd.types import MANIFOLD_DTYPE
    lib = bounce_amd.load_library()
    t = time.perf_counter()
    pool, pairs, xfs = workload(lib, args.pairs, args.pool)
    log("workload: %d pairs over %d hulls in %.1f s" % (len(pairs), pool.count, time.perf_counter() - t))

    cpu = None
    if not args.no_cpu:  # before any GPU initialisation: the baseline forks workers
        t = time.perf_counter()
        cpu = cpu_baseline(pool, pairs, xfs, args.cpu_sample)
        log("cpu baseline: %.0f pairs/s (%s) in %.1f s" % (cpu["value"], cpu["kind"], time.perf_counter() - t))
    mesh_world = mesh_cpu = None
    if not args.no_mesh:
        from bounce_amd.workloads import gen_mesh_world
        mesh_world = gen_mesh_world(lib, 7)
        if not args.no_cpu:  # also before the GPU is initialised
            mesh_cpu = mesh_cpu_baseline(*mesh_world, 100000)
            log("mesh cpu baseline: %.0f contacts/s (%s)" % (mesh_cpu["value"], mesh_cpu["kind"]))
    convex_world = convex_cpu = None
    if not args.no_convex:
        from bounce_amd.workloads import gen_convex_pairs
        convex_world = gen_convex_pairs(lib, 5, CONVEX_PAIRS)
        if not args.no_cpu:  # also before the GPU is initialised
            convex_cpu = cpu_baseline(convex_world[0], convex_world[2], convex_world[3], CONVEX_PAIRS,
                                      shapes=convex_world[1])
            log("convex cpu baseline: %.0f pairs/s (%s)" % (convex_cpu["value"], convex_cpu["kind"]))
    warm_world = warm_cpu = None
    if not args.no_warm:
        from bounce_amd.workloads import gen_warm_world
        t = time.perf_counter()
        warm_world = gen_warm_world(lib, 11)
        log("warm world: %d bodies, %d contacts in %.1f s"
            % (len(warm_world["body_shapes"]), len(warm_world["contacts"]), time.perf_counter() - t))
        if not args.no_cpu:  # also before the GPU is initialised
            warm_cpu = warm_cpu_baseline(warm_world, 4)
            log("warm world cpu baseline: %.0f contacts/s warm (%s)" % (warm_cpu[0]["value"], warm_cpu[0]["kind"]))
    rain_world = rain_cpu = None
    if not args.no_rain:
        from bounce_amd.workloads import gen_rain_world, gen_warm_world
        t = time.perf_counter()
        rain_world = gen_rain_world(lib, 11, base=warm_world if warm_world is not None else gen_warm_world(lib, 11))
        log("rain world: %d bodies in %.1f s" % (len(rain_world["body_shapes"]), time.perf_counter() - t))
        if not args.no_cpu:  # also before the GPU is initialised
            t = time.perf_counter()
            rain_cpu = rain_cpu_baseline(rain_world, args.rain_steps)
            log("rain world cpu baseline: %.0f contacts/s (%s) in %.1f s"
                % (rain_cpu["value"], rain_cpu["kind"], time.perf_counter() - t))
    work = algorithmic_work(pool, pairs, xfs, max(1, len(pairs) // 8192))
    log("algorithmic work: %.3g SAT flops, %.3g SAT bytes per batch" % (work["sat_flops"], work["sat_bytes"]))

    n = len(pairs)
    with bounce_amd.Context(0) as ctx:
        ctx.upload_hulls(pool)
        ctx.upload_shapes(pool.shapes())
        dp, dx, do = ctx.alloc(pairs.nbytes), ctx.alloc(xfs.nbytes), ctx.alloc(n * MANIFOLD_DTYPE.itemsize)
        dp.upload(pairs)
        dx.upload(xfs)
        if args.copy_calibration:
            a, b = ctx.alloc(args.copy_calibration), ctx.alloc(args.copy_calibration)
            ctx.copy_on_device(b, a, args.copy_calibration)
            a.free()
            b.free()
        if args.tuning == "auto":
            t = time.perf_counter()
            shape, tms = ctx.autotune(n, dp, dx, do, repeats=3)
            log("autotune: waves/SIMD %d, tail chunk %d (%.3f ms per batch) in %.1f s"
                % (shape[0], shape[1], tms, time.perf_counter() - t))
        else:
            w, k = (int(x) for x in args.tuning.split(","))
            ctx.set_sat_tuning(w, k)
        tuning = ctx.sat_tuning()
        for _ in range(args.warmup):
            ctx.collide_device(n, dp, dx, do)
        ctx.synchronize()
        ctx.reset_kernel_stats()
        ctx.set_profiling(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.collide_device(n, dp, dx, do)
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.set_profiling(False)
        stats = {k: ctx.kernel_stats(k) for k in ("cache_hull", "sat_hull", "contact_hull")}
        out = do.download(np.zeros(n, MANIFOLD_DTYPE))
        forms = other_forms(ctx, n, dp, dx, do, args.steps, out)
        # The plain drop-in form: one b3amd_collide per step on host arrays (PCIe both ways).
        plain_ms = None
        if not args.no_plain:
            ctx.collide(pairs, xfs)
            t2 = time.perf_counter()
            plain_steps = 3
            for _ in range(plain_steps):
                ctx.collide(pairs, xfs)
            plain_ms = (time.perf_counter() - t2) * 1000.0 / plain_steps
    ms = (t1 - t0) * 1000.0 / args.steps
    value = n / (ms / 1000.0)
    launches, sat_total = stats["sat_hull"]
    sat_ms = sat_total / max(1, launches)
    achieved_tflops = work["sat_flops"] / (sat_ms / 1000.0) / 1e12
    achieved_gbs = work["sat_bytes"] / (sat_ms / 1000.0) / 1e9
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched hull-hull narrowphase: %d pairs over %d hulls of 8-64 points on ellipsoids "
                        "(semi-axes U[0.25,2] m), uniform orientations, SAT signed gap ~ N(-0.0127, 0.05) m, "
                        "no feature caches (every pair through the SAT)" % (n, pool.count),
            "pairs": n,
            "hull_pool": pool.count,
            "overlap_fraction": float((out["pointCount"] > 0).mean()),
            "parallelism": "single GPU",
            "sat_launch_shape": {"waves_per_simd": tuning[0], "tail_chunk": tuning[1],
                                 "chosen_by": "autotune" if args.tuning == "auto" else "flag"},
        },
        "kernels_ms": {k: (v[1] / max(1, v[0])) for k, v in stats.items()},
        "roofline": {
            "kernel": "sat_hull",
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / VALU_PEAK_TFLOPS,
            "frac_unfused_peak": achieved_tflops / VALU_UNFUSED_TFLOPS,
            "hbm_achieved_GBs": achieved_gbs,
            "hbm_frac": achieved_gbs / HBM_PEAK_GBS,
            "flops_per_launch": work["sat_flops"],
            "bytes_per_launch": work["sat_bytes"],
            "traffic": measured_traffic(n),
            "traffic_source": "rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (scaled by a 1 GiB copy), "
                              "profiles/sat_hull_1M_pairs_hbm_traffic.md",
        },
        "cpu_baseline": cpu,
        "plain_drop_in": None if plain_ms is None else {
            "value": n / (plain_ms / 1000.0), "unit": "pairs/s", "ms_per_step": plain_ms,
            "what": "b3amd_collide on host arrays, one call per step: H2D of pairs and transforms, the "
                    "same kernels, D2H of manifolds"},
        "separation_error_fp64": separation_error_fp64(out, xfs),
        "other_forms": forms,
    }
    if not args.no_mesh:
        result["mesh"] = mesh_run(mesh_world, 3, mesh_cpu)
        log("mesh: %.0f contacts/s resident (%.3f ms per batch), %.0f contacts/s from host arrays"
            % (result["mesh"]["contacts_per_s"], result["mesh"]["ms_per_batch"],
               result["mesh"]["host_arrays_contacts_per_s"]))
    if not args.no_convex:
        result["convex"] = convex_run(convex_world, 5, convex_cpu)
        log("convex: %.0f pairs/s resident (%.3f ms per batch, convex kernel %.3f ms, %.3f of VALU peak)"
            % (result["convex"]["pairs_per_s"], result["convex"]["ms_per_batch"],
               result["convex"]["convex_kernel_ms"], result["convex"]["roofline"]["frac"]))
    if not args.no_pyramid:
        result["pyramid"] = pyramid_run(lib)
        log("pyramid: %d/%d steps equal to the reference, %.3f ms per step; whole step on the library: %d/%d "
            "steps equal, %.3f ms per step"
            % (result["pyramid"]["steps_equal_to_reference"], result["pyramid"]["steps"],
               result["pyramid"]["ms_per_step"], result["pyramid"]["whole_step"]["steps_equal_to_reference"],
               result["pyramid"]["whole_step"]["steps"], result["pyramid"]["whole_step"]["ms_per_step"]))
    if not args.no_warm:
        result["warm_world"] = warm_world_run(warm_world, args.warm_steps, warm_cpu)
        log("warm world: %.0f contacts/s over %d warm steps (%.3f ms per step), cold step %.3f ms, "
            "cache hits %.3f" % (result["warm_world"]["warm_contacts_per_s"], args.warm_steps - 1,
                                 result["warm_world"]["warm_ms_per_step"], result["warm_world"]["cold_step_ms"],
                                 result["warm_world"]["cache_hit_fraction"]))
    if not args.no_serving:
        result["serving"] = serving_run(lib, args.serving_worlds, args.serving_rate, args.serving_seconds)
    if not args.no_rain:
        t = time.perf_counter()
        result["rain_world"] = rain_world_run(lib, rain_world, args.rain_steps, rain_cpu)
        r = result["rain_world"]
        log("rain world: %.0f contacts/s, %.3f ms per step, %d created, %d destroyed, %d contacts at the end, "
            "parity %s in %.1f s" % (r["contacts_per_s"], r["ms_per_step"], r["contacts_created"],
                                     r["contacts_destroyed"], r["contacts_at_end"], r.get("sampled_parity"),
                                     time.perf_counter() - t))
    print(json.dumps(result))


if __name__ == "__main__":
    main()
