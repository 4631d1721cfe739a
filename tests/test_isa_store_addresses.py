"""Guard against the GPU memory fault class of divergent merged stores: every global, flat
or buffer access in the AMD assembly of the cache, contact, serving and contact-setup
kernels must address through VGPRs written in every active lane (tests/isa_check.py states
the analysis).  The checker itself is pinned on synthetic shapes and on the cache/contact
kernels of the revision that faulted on the GPU (its merged stores are flagged, the fixed
kernels are not).

Scope: narrowphase_hull.hip (the cache and contact kernels where the fault occurred),
serving.hip and contact_init.hip.  The checker follows uniform branches on both edges, so
kernels that select a path by a uniform flag and redefine an address on each path
(narrowphase_convex.hip's optional sweeps) would be reported falsely, and the SAT and mesh
kernels' loop nests exceed its fixpoint budget; those files are not checked here.  In
narrowphase_hull.hip, contact_wide_kernel is left out for the same reason (its loop over the
listed pairs around four inlined clip loops); it runs on one lane per block, every other lane
leaving before the first memory access, so no access of it can merge lanes that defined the
address differently."""
import concurrent.futures
import os
import shutil
import subprocess
import tempfile

import pytest

import isa_check

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CHECKED = ["narrowphase_hull.hip", "serving.hip", "contact_init.hip"]
FAULTING_REVISION = "7ab00eb^"  # the cache/contact kernels before their stores were re-addressed


def device_asm(src, root, out):
    from bounce_amd.build import COMMON
    flags = [f for f in COMMON if not f.startswith("-I") and f != "-fPIC"]
    cmd = [HIPCC] + flags + ["-I" + os.path.join(root, "include"), "-I" + os.path.join(root, "bounce_amd", "csrc"),
                             "-x", "hip", "--cuda-device-only", "-S", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-2000:]))
    return open(out).read()


@pytest.fixture(scope="module")
def asm():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    tmp = tempfile.mkdtemp(prefix="b3amd_isa_")
    try:
        with concurrent.futures.ThreadPoolExecutor(len(CHECKED)) as ex:
            texts = list(ex.map(lambda f: device_asm(os.path.join(ROOT, "bounce_amd", "csrc", f), ROOT,
                                                     os.path.join(tmp, f + ".s")), CHECKED))
        yield dict(zip(CHECKED, texts))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


SINGLE_LANE_KERNELS = ("contact_wide_kernel",)


@pytest.mark.parametrize("src", CHECKED)
def test_kernels_address_only_written_registers(asm, src):
    findings = isa_check.check_assembly(asm[src], skip=SINGLE_LANE_KERNELS)
    assert not findings, "%s: accesses through VGPRs unwritten in some active lane: %s" % (
        src, {k: v[:4] for k, v in findings.items()})


def test_checker_flags_the_faulting_revision():
    """The cache/contact kernels as they were when a mixed-shape batch faulted: the checker
    reports the merged stores whose address sphere/capsule lanes never wrote."""
    if not os.path.exists(HIPCC) or not os.path.isdir(os.path.join(ROOT, ".git")):
        pytest.skip("needs hipcc and the repository history")
    tmp = tempfile.mkdtemp(prefix="b3amd_isa_old_")
    try:
        arch = subprocess.run(["git", "-C", ROOT, "archive", FAULTING_REVISION, "bounce_amd/csrc", "include"],
                              capture_output=True)
        if arch.returncode != 0:
            pytest.skip("revision %s not in this history" % FAULTING_REVISION)
        subprocess.run(["tar", "-x", "-C", tmp], input=arch.stdout, check=True)
        text = device_asm(os.path.join(tmp, "bounce_amd", "csrc", "narrowphase_hull.hip"), tmp,
                          os.path.join(tmp, "old.s"))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    findings = isa_check.check_assembly(text)
    assert any("contact_hull_kernel" in k for k in findings), findings.keys()
    assert all(op.startswith("global_store") for k in findings for op, _ in findings[k])


def kernel(body):
    return "k:\n" + "\n".join("\t" + line for line in body.strip().splitlines()) + "\n.Lfunc_end0:\n"


def test_divergent_definition_then_wider_store_is_flagged():
    text = kernel("""
v_cmp_gt_u32_e32 vcc, 3, v0
s_and_saveexec_b64 s[2:3], vcc
s_cbranch_execz .LBB0_2
v_mad_u64_u32 v[2:3], s[4:5], v0, 4, s[6:7]
.LBB0_2:
s_or_b64 exec, exec, s[2:3]
global_store_dword v[2:3], v1, off
s_endpgm""")
    assert isa_check.check_assembly(text)["k"] == [("global_store_dword v[2:3] v1 off", 2),
                                                  ("global_store_dword v[2:3] v1 off", 3)]


def test_definition_before_divergence_passes():
    text = kernel("""
v_mad_u64_u32 v[2:3], s[4:5], v0, 4, s[6:7]
v_cmp_gt_u32_e32 vcc, 3, v0
s_and_saveexec_b64 s[2:3], vcc
s_cbranch_execz .LBB0_2
v_mov_b32_e32 v1, 7
global_store_dword v[2:3], v1, off
.LBB0_2:
s_or_b64 exec, exec, s[2:3]
global_store_dword v[2:3], v0, off
s_endpgm""")
    assert isa_check.check_assembly(text) == {}


def test_if_else_both_defining_passes_and_one_sided_is_flagged():
    both = kernel("""
v_cmp_gt_u32_e32 vcc, 3, v0
s_and_saveexec_b64 s[2:3], vcc
s_xor_b64 s[2:3], exec, s[2:3]
s_cbranch_execz .LBB0_2
v_mad_u64_u32 v[2:3], s[4:5], v0, 4, s[6:7]
.LBB0_2:
s_or_saveexec_b64 s[2:3], s[2:3]
s_xor_b64 exec, exec, s[2:3]
s_cbranch_execz .LBB0_4
v_mad_u64_u32 v[2:3], s[4:5], v0, 8, s[8:9]
.LBB0_4:
s_or_b64 exec, exec, s[2:3]
global_store_dword v[2:3], v1, off
s_endpgm""")
    assert isa_check.check_assembly(both) == {}
    one = both.replace("v_mad_u64_u32 v[2:3], s[4:5], v0, 8, s[8:9]", "v_mov_b32_e32 v5, 0")
    assert "k" in isa_check.check_assembly(one)


def test_uniform_skip_of_an_empty_region_is_vacuous():
    """A region skipped because no lane is active defines nothing but needs nothing."""
    text = kernel("""
v_cmp_gt_u32_e32 vcc, 3, v0
s_and_saveexec_b64 s[2:3], vcc
s_cbranch_execz .LBB0_2
v_mad_u64_u32 v[2:3], s[4:5], v0, 4, s[6:7]
global_load_dword v4, v[2:3], off
global_store_dword v[2:3], v4, off
.LBB0_2:
s_or_b64 exec, exec, s[2:3]
s_endpgm""")
    assert isa_check.check_assembly(text) == {}
