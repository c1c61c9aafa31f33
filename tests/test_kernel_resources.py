"""Every kernel of the HIP sources compiles for gfx950 without scratch
accesses and without register spills: a kernel that indexes an array of its
argument struct at run time, or runs out of VGPRs, falls back to scratch,
which runs several times slower.  Device-only assembly (no GPU needed); the
kernel descriptors' metadata give the sizes, the kernel bodies the
instructions.  A private segment with no scratch instruction in the body is
the register scavenger's reserved slot (kept when a kernel uses every SGPR)
and costs nothing."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import REPO

CSRC = os.path.join(REPO, "matjes_amd", "csrc")


def _device_asm(src):
    from matjes_amd import build
    flags = [f for f in build.HIPCC_FLAGS if f not in ("-shared", "-fPIC")] + build.SOURCE_FLAGS.get(src, [])
    cmd = [build.hipcc()] + flags + ["--cuda-device-only", "-S", "-I", CSRC, "-I", os.path.join(REPO, "include"),
                                     "-o", "-", os.path.join(CSRC, src)]
    return subprocess.run(cmd, capture_output=True, text=True, check=True).stdout


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not present")
@pytest.mark.parametrize("src", ["mj_kernels.hip", "mj_batch.hip", "mj_adjoint.hip", "mj_mc.hip", "mj_fft.hip"])
def test_kernels_use_no_scratch(src):
    asm = _device_asm(src)
    names = re.findall(r"^\s+\.name:\s+(\S+)", asm, re.M)
    priv = [int(v) for v in re.findall(r"^\s+\.private_segment_fixed_size:\s+(\d+)", asm, re.M)]
    spill = [int(v) for v in re.findall(r"^\s+\.vgpr_spill_count:\s+(\d+)", asm, re.M)]
    assert names and len(names) == len(priv) == len(spill), src
    bodies = _bodies(asm)
    def scratch_ops(n):
        assert n in bodies, n
        return _scratch_ops(bodies[n])
    bad = [(n, p, s, scratch_ops(n)) for n, p, s in zip(names, priv, spill) if s or (p and scratch_ops(n))]
    assert not bad, bad


def _bodies(asm):
    """kernel name -> its instructions (label to .Lfunc_end)"""
    return {m.group(1): m.group(2) for m in re.finditer(r"^(\S+):\s+;\s+@\S+\n(.*?)^\.Lfunc_end", asm, re.M | re.S)}


def _scratch_ops(body):
    return len(re.findall(r"^\s+(scratch_\w+|buffer_(?:load|store)\w*)\s", body, re.M))


def test_scratch_parser_sees_planted_accesses():
    """The parser finds a scratch store and a MUBUF scratch load planted in a
    kernel body, and nothing in a body without them."""
    asm = ("k1:                                     ; @k1\n\tv_mov_b32 v0, 0\n\tscratch_store_dword off, v0, s0 offset:4\n"
           ".Lfunc_end0:\n"
           "k2:                                     ; @k2\n\tbuffer_load_dword v1, off, s[0:3], 0 offset:8\n.Lfunc_end1:\n"
           "k3:                                     ; @k3\n\tv_add_f64 v[0:1], v[2:3], v[4:5]\n.Lfunc_end2:\n")
    b = _bodies(asm)
    assert set(b) == {"k1", "k2", "k3"}
    assert _scratch_ops(b["k1"]) == 1 and _scratch_ops(b["k2"]) == 1 and _scratch_ops(b["k3"]) == 0
