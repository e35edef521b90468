"""The mark queue's room check as the compiler builds it (CPU only: hipcc
cross-compiles gfx950).

A light claims `extra` entries of its mark sub-queue (seg entries) at `base`
and queues min(extra, room) chunks, room = seg - base saturated at zero.
Written as fit = base >= seg ? 0 : min(extra, seg - base), -O3 marked the
difference no-wrap (sub nuw) and folded the select away, so an overflowing
light took fit = extra and wrote past its segment; this test fails on that
form (no llvm.usub.sat, a sub nuw in its place).  The source now states the room
as a saturating subtraction; this test compiles hl_cull.hip to LLVM IR with
the library's own flags and checks that every mark kernel still computes it
with llvm.usub.sat, so a rewrite or a compiler change that brings the fold
back fails here, not on the device.
"""
import re
import shutil
import subprocess

import pytest

from helio_amd import build

MARK_KERNELS = ("k_cluster_mark", "k_batch_mark")


@pytest.fixture(scope="module")
def cull_ir(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not shutil.which(hipcc):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("ir") / "hl_cull.ll"
    cmd = [hipcc, *build.HIPCC_FLAGS, "--cuda-device-only", "-S", "-emit-llvm",
           str(build.CSRC / "hl_cull.hip"), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return out.read_text()


def kernel_bodies(ir, name):
    """The IR bodies of the kernels whose mangled name holds `name` followed by
    its parameter list (k_cluster_mark, not k_cluster_mark_tail)."""
    pat = re.compile(r"^define [^\n]*@(_Z[^\s(]*" + name + r"E[^\s(]*)\(.*?^}", re.M | re.S)
    return [m.group(0) for m in pat.finditer(ir) if "tail" not in m.group(1)]


@pytest.mark.parametrize("name", MARK_KERNELS)
def test_mark_kernel_room_is_a_saturating_subtraction(cull_ir, name):
    bodies = kernel_bodies(cull_ir, name)
    assert len(bodies) == 1, f"{name}: {len(bodies)} definitions"
    body = bodies[0]
    # the claim on the sub-queue counter, then the room from the returned base
    assert "atomicrmw add" in body
    assert "@llvm.usub.sat.i32" in body, f"{name}: the sub-queue room is not a saturating subtraction"
