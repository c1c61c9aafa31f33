#!/usr/bin/env python3
"""Build the reference Fortran (Matjes) from its own sources into oracle/_ref/.

Test infrastructure only: the objects and the driver built here are the
checker that pins the oracle and writes golden fixtures; nothing in the
product links or loads them.

Recipe (no CMake, no stand-ins): the Fortran file list is read from the
reference's own per-directory lists (src/*/CMakeLists.txt, the `main_f90`
sets), the module dependency order is derived from `module`/`use`
statements, and every file is compiled with ROCm's amdflang using the
preprocessor switches of the reference's script configuration
(cmake/configlist/config.cmake_manjaro: CPP_SCRIPT) and no optional library
(no FFTW/MKL/Eigen/BLAS/MPI/NetCDF, so the dense backend `Hamiltonian_mode 5`
is the only matrix backend and the FFT Hamiltonians are unavailable).
The reference's main program is NOT linked (it pulls the tight-binding code,
which needs an absent BLAS); instead oracle/ref/mjref_driver.f90 calls the
reference's own setup_simu / get_eff_field / energy / spindynamics.
Objects are LLVM bitcode (-flto) in an archive, so the link keeps only what
the driver reaches; unreachable procedures (e.g. the GSL-backed Laguerre
wrapper that shares an object with reachable code) are dropped rather than
resolved, and no stand-in symbol is supplied for any missing library.

Usage: python oracle/ref/build_ref.py [--ref /root/reference] [-j 8]
"""
import argparse
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(REPO, "oracle", "_ref")

INTRINSIC = {"iso_c_binding", "iso_fortran_env", "omp_lib", "ieee_arithmetic",
             "ieee_exceptions", "ieee_features", "omp_lib_kinds", "mpi", "mpi_f08"}
FLAGS = ["-cpp", "-O2", "-fopenmp", "-DCPP_SCRIPT", "-flto"]


def fortran_files(ref):
    files = []
    top = open(os.path.join(ref, "src", "CMakeLists.txt")).read()
    for inc in re.findall(r"include\(\s*(src/[^)\s]+/CMakeLists.txt)\s*\)", top):
        cm = os.path.join(ref, inc)
        if not os.path.isfile(cm):
            continue
        text = open(cm).read()
        for block in re.findall(r"set\(\s*main_f90\s+\$\{main_f90\}(.*?)\)", text, re.S):
            for tok in block.split():
                if tok.lower().endswith((".f90", ".f")):
                    files.append(os.path.join(ref, tok))
    files.append(os.path.join(ref, "src", "stdtypes.f90"))
    seen, out = set(), []
    for f in files:
        if f not in seen and os.path.isfile(f):
            seen.add(f)
            out.append(f)
    return out


def scan(path):
    mods, uses = set(), set()
    for line in open(path, errors="replace"):
        s = line.split("!")[0].strip().lower()
        m = re.match(r"module\s+(\w+)\s*$", s)
        if m and m.group(1) != "procedure":
            mods.add(m.group(1))
        m = re.match(r"use\s*(?:,\s*\w+\s*)?(?:::)?\s*(\w+)", s)
        if m:
            uses.add(m.group(1))
    return mods, uses


def order(files):
    info = {f: scan(f) for f in files}
    owner = {}
    for f, (mods, _) in info.items():
        for m in mods:
            owner[m] = f
    deps = {}
    for f, (mods, uses) in info.items():
        deps[f] = {owner[u] for u in uses if u in owner and owner[u] != f and u not in INTRINSIC}
    levels, done = [], set()
    remaining = set(files)
    while remaining:
        lvl = sorted(f for f in remaining if deps[f] <= done)
        if not lvl:
            raise RuntimeError("module dependency cycle among: %s" % sorted(remaining)[:5])
        levels.append(lvl)
        done |= set(lvl)
        remaining -= set(lvl)
    return levels


def compile_one(f, fc, moddir, objdir):
    obj = os.path.join(objdir, os.path.relpath(f, "/").replace("/", "_") + ".o")
    cmd = [fc] + FLAGS + ["-module-dir", moddir, "-I", moddir, "-c", f, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return f, obj, r.returncode, r.stderr[-2000:]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--fc", default="amdflang")
    a = ap.parse_args()
    if not os.path.isdir(os.path.join(a.ref, "src")):
        sys.exit("reference sources not found at %s" % a.ref)
    moddir = os.path.join(OUT, "mod")
    objdir = os.path.join(OUT, "obj")
    os.makedirs(moddir, exist_ok=True)
    os.makedirs(objdir, exist_ok=True)
    levels = order(fortran_files(a.ref))
    objs, failed = [], []
    with ThreadPoolExecutor(a.j) as ex:
        for lvl in levels:
            for f, obj, rc, err in ex.map(lambda f: compile_one(f, a.fc, moddir, objdir), lvl):
                if rc:
                    failed.append((f, err))
                else:
                    objs.append(obj)
            if failed:
                for f, err in failed:
                    print("FAILED", f, "\n", err, file=sys.stderr)
                sys.exit(1)
    lib = os.path.join(OUT, "libmatjes_ref.a")
    if os.path.exists(lib):
        os.remove(lib)
    subprocess.check_call(["ar", "rcs", lib] + objs)
    drv = os.path.join(HERE, "mjref_driver.f90")
    exe = os.path.join(OUT, "mjref_driver")
    subprocess.check_call([a.fc] + FLAGS + ["-module-dir", moddir, "-I", moddir, drv,
                           lib, "-o", exe])
    print("built", exe)
    # the t_H_hip and fft_H_hip backends (matjes_amd/fortran/ham_hip.f90,
    # fft_ham_hip.f90) inside the reference's class tree, driven by
    # oracle/ref/mjref_hipham.f90; needs libmatjes_hip.so
    hiplib = os.path.join(REPO, "matjes_amd", "libmatjes_hip.so")
    if os.path.isfile(hiplib):
        fdir = os.path.join(REPO, "matjes_amd", "fortran")
        objs2 = []
        for src in ("m_matjes_hip.f90", "ham_hip.f90", "fft_ham_hip.f90"):
            obj = os.path.join(objdir, src.replace(".f90", ".o"))
            subprocess.check_call([a.fc] + FLAGS + ["-module-dir", moddir, "-I", moddir, "-c",
                                   os.path.join(fdir, src), "-o", obj])
            objs2.append(obj)
        exe2 = os.path.join(OUT, "mjref_hipham")
        subprocess.check_call([a.fc] + FLAGS + ["-module-dir", moddir, "-I", moddir,
                               os.path.join(HERE, "mjref_hipham.f90")] + objs2 +
                              [lib, "-L" + os.path.dirname(hiplib), "-lmatjes_hip",
                               "-Wl,-rpath,$ORIGIN/../../matjes_amd", "-o", exe2])
        print("built", exe2)


if __name__ == "__main__":
    main()
