"""Builds libbounce_amd.so in-tree for gfx950 with hipcc.

Flags that the bit-exact contract depends on:
  -ffp-contract=off          no FMA contraction (the reference is x86-64 SSE2, no FMA)
  IEEE fp32 division/sqrt    hipcc's default correctly rounded fp32 div/sqrt is kept
                             (no -ffast-math, no -fno-hip-fp32-correctly-rounded-divide-sqrt)
  no denormal flushing       no -fgpu-flush-denormals-to-zero
  -Xclang -target-feature -Xclang -packed-fp32-ops   packed FP32 instructions disabled
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbounce_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["narrowphase_hull.hip", "sat_persistent.hip", "serving.hip", "narrowphase_convex.hip", "narrowphase_mesh.hip", "contact_init.hip", "broadphase.hip", "contact_manager.hip", "host_world.cpp", "api.cpp", "workload.cpp",
           "static_tree.cpp", "quickhull.cpp", "hullfile.cpp"]

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wall", "-Wno-unused-function",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose=False):
    """Test infrastructure: the oracle restatement and, where the reference tree is
    present, the reference compiled in place (oracle/_ref)."""
    cmd = ["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j8", "all"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s\n%s" % (r.stdout, r.stderr))
    if verbose and r.stdout.strip():
        print(r.stdout.strip())


if __name__ == "__main__":
    build_library(verbose=True)
    if "--oracle" in sys.argv:
        build_oracle(verbose=True)
