"""Every kernel of the HIP sources compiles for gfx950 without private
(scratch) memory and without register spills: a kernel that indexes an array
of its argument struct at run time, or runs out of VGPRs, falls back to
scratch, which runs several times slower.  Device-only assembly (no GPU
needed); the kernel descriptors' metadata give the sizes."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import REPO

CSRC = os.path.join(REPO, "matjes_amd", "csrc")


def _device_asm(src):
    from matjes_amd import build
    flags = [f for f in build.HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
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
    bad = [(n, p, s) for n, p, s in zip(names, priv, spill) if p or s]
    assert not bad, bad
