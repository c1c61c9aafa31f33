"""Build the native pieces in-tree.

libmatjes_hip.so  -- the HIP library (gfx950) behind include/matjes_hip.h
libmjf.a / mjf_llg -- the Fortran interface module and the Fortran host
                      driver (amdflang), when amdflang is present
"""
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libmatjes_hip.so")
FDIR = os.path.join(PKG, "fortran")

HIP_SOURCES = ["mj_api.hip", "mj_kernels.hip", "mj_fft.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # bit-for-bit reference arithmetic in the per-site algebra: no contraction
    # (the stencil sum uses explicit fma)
    "-ffp-contract=off",
    # packed FP32 instructions stay disabled on these boxes
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_hip(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(REPO, "include")] + srcs + \
        ["-L/opt/rocm/lib", "-lrocfft", "-Wl,-rpath,/opt/rocm/lib", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout[-4000:] + r.stderr[-8000:])
    if verbose and r.stderr:
        print(r.stderr[-4000:])
    os.replace(tmp, LIB)
    return LIB


def flang():
    for c in ("/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran():
    """Compile the iso_c_binding module and the Fortran host driver against the .so."""
    fc = flang()
    if fc is None:
        return None
    mod = os.path.join(FDIR, "m_matjes_hip.f90")
    drv = os.path.join(FDIR, "mjf_llg.f90")
    exe = os.path.join(FDIR, "mjf_llg")
    bdir = os.path.join(FDIR, "_build")
    os.makedirs(bdir, exist_ok=True)
    cmd = [fc, "-O2", "-module-dir", bdir, "-I", bdir, mod, drv, "-L" + PKG, "-lmatjes_hip",
           "-Wl,-rpath," + PKG, "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("amdflang failed:\n" + r.stdout[-4000:] + r.stderr[-4000:])
    return exe


def build_all(verbose=False):
    build_hip(verbose)
    if os.path.exists(os.path.join(FDIR, "mjf_llg.f90")):
        build_fortran()


if __name__ == "__main__":
    build_all(verbose=True)
    print("built", LIB)
