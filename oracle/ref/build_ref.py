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

With matjes_amd/libmatjes_hip.so built, a second archive adds the HIP
backends to the reference's own factories: copies of ham_public.f90 and
fft_ham_public.f90 (written under oracle/_ref/hipmode/src, never committed)
gain one `case` each in set_Ham_mode_io / get_Htype / get_Htype_N
(ham_public.f90:21-162) and set_fft_H_mode_io / get_fft_H_single /
get_fft_H_N (fft_ham_public.f90:15-120), selected by the reference's own
input keys `Hamiltonian_mode 8` (t_H_hip) and `Hamiltonian_fft_mode 4`
(fft_H_hip).  Only the files that use those two modules, directly or not,
are recompiled against the copies; oracle/ref/mjref_hipham.f90 links that
archive, so setup_simu itself builds every term on the device backend.

Usage: python oracle/ref/build_ref.py [--ref /root/reference] [-j 8]
"""
import argparse
import os
import re
import shutil
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


def obj_name(f, objdir):
    return os.path.join(objdir, os.path.relpath(f, "/").replace("/", "_") + ".o")


def compile_one(f, fc, moddir, objdir):
    obj = obj_name(f, objdir)
    cmd = [fc] + FLAGS + ["-module-dir", moddir, "-I", moddir, "-c", f, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return f, obj, r.returncode, r.stderr[-2000:]


# the modes the patched factories add: the next free values of the
# reference's `Hamiltonian_mode` (1-7 taken) and `Hamiltonian_fft_mode` (0-3)
HIP_HAM_MODE = 8
HIP_FFT_MODE = 4

# (file, use line to add, line it goes after, {subroutine: (anchor, lines)}):
# each block is inserted before the first line of that subroutine matching
# its anchor, i.e. as one more `case` ahead of the error cases
FACTORIES = [
    ("src/Hamiltonian_type/ham_public.f90", "use m_H_hip, only: t_H_hip", r"\s*use\s+m_H_dense\s*$", {
        "set_ham_mode_io": (r"\s*case\s+default", [
            "        case(%d)" % HIP_HAM_MODE,
            "            write(*,'(/A/)') \"Using Hamiltonian implementation: HIP (matjes_amd)\""]),
        "get_htype": (r"\s*case\s*\(\s*-1\s*\)", [
            "        case(%d)" % HIP_HAM_MODE,
            "            allocate(t_H_hip::H_out)"]),
        "get_htype_n": (r"\s*case\s*\(\s*-1\s*\)", [
            "        case(%d)" % HIP_HAM_MODE,
            "            allocate(t_H_hip::H_out(N))"]),
    }),
    ("src/Hamiltonian_type/fft_ham_public.f90", "use m_fft_H_hip, only: fft_H_hip", r"\s*use\s+m_fft_H_fftwmpi\b", {
        "set_fft_h_mode_io": (r"\s*case\s*\(\s*0\s*\)", [
            "        case(%d)" % HIP_FFT_MODE,
            "            write(*,'(/A/)') \"Using Hamiltonian_fft implementation: HIP (matjes_amd)\""]),
        "get_fft_h_single": (r"\s*case\s*\(\s*0\s*\)", [
            "        case(%d)" % HIP_FFT_MODE,
            "            allocate(fft_H_hip::H_out)"]),
        "get_fft_h_n": (r"\s*case\s*\(\s*0\s*\)", [
            "        case(%d)" % HIP_FFT_MODE,
            "            allocate(fft_H_hip::H_out(N))"]),
    }),
]


def patch_factory(src, dst, use_line, use_after, cases):
    """Write `src` to `dst` with the use line and the `case` blocks added;
    fails when an anchor is not found exactly where expected."""
    out, sub, done, used = [], None, set(), False
    for line in open(src, errors="replace").read().split("\n"):
        m = re.match(r"\s*subroutine\s+(\w+)", line, re.I)
        if m:
            sub = m.group(1).lower()
        if sub in cases and sub not in done and re.match(cases[sub][0], line, re.I):
            out.extend(cases[sub][1])
            done.add(sub)
        out.append(line)
        if not used and re.match(use_after, line, re.I):
            out.append(use_line)
            used = True
    if not used or done != set(cases):
        raise RuntimeError("%s: anchors not found (use: %s, cases: %s)" % (src, used, sorted(set(cases) - done)))
    with open(dst, "w") as f:
        f.write("\n".join(out))


def build_hip_mode(ref, files, fc, moddir, objdir, jobs):
    """The archive whose factories also allocate t_H_hip / fft_H_hip; returns
    its path and the module directory of that build."""
    hout = os.path.join(OUT, "hipmode")
    hsrc, hmod, hobj = (os.path.join(hout, d) for d in ("src", "mod", "obj"))
    for d in (hsrc, hmod, hobj):
        os.makedirs(d, exist_ok=True)
    for m in os.listdir(moddir):
        shutil.copy2(os.path.join(moddir, m), os.path.join(hmod, m))
    swap = {}
    for rel, use_line, use_after, cases in FACTORIES:
        dst = os.path.join(hsrc, os.path.basename(rel))
        patch_factory(os.path.join(ref, rel), dst, use_line, use_after, cases)
        swap[os.path.join(ref, rel)] = dst
    fdir = os.path.join(REPO, "matjes_amd", "fortran")
    extra = [os.path.join(fdir, f) for f in ("m_matjes_hip.f90", "ham_hip.f90", "fft_ham_hip.f90")]
    hfiles = [swap.get(f, f) for f in files] + extra
    info = {f: scan(f) for f in hfiles}
    # every file that uses a module of a changed file, transitively
    changed = set(swap.values()) | set(extra)
    grow = True
    while grow:
        mods = set().union(*(info[f][0] for f in changed))
        more = {f for f in hfiles if f not in changed and info[f][1] & mods}
        changed |= more
        grow = bool(more)
    objs, failed = [], []
    with ThreadPoolExecutor(jobs) as ex:
        for lvl in order(sorted(changed)):
            for f, obj, rc, err in ex.map(lambda f: compile_one(f, fc, hmod, hobj), lvl):
                if rc:
                    failed.append((f, err))
                else:
                    objs.append(obj)
            if failed:
                for f, err in failed:
                    print("FAILED", f, "\n", err, file=sys.stderr)
                sys.exit(1)
    objs += [obj_name(f, objdir) for f in hfiles if f not in changed]
    lib = os.path.join(hout, "libmatjes_ref_hip.a")
    if os.path.exists(lib):
        os.remove(lib)
    subprocess.check_call(["ar", "rcs", lib] + objs)
    print("recompiled %d of %d files for the HIP factory modes" % (len(changed), len(hfiles)))
    return lib, hmod


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
    # fft_ham_hip.f90) inside the reference's class tree and its factories,
    # driven by oracle/ref/mjref_hipham.f90; needs libmatjes_hip.so
    hiplib = os.path.join(REPO, "matjes_amd", "libmatjes_hip.so")
    if os.path.isfile(hiplib):
        hlib, hmod = build_hip_mode(a.ref, fortran_files(a.ref), a.fc, moddir, objdir, a.j)
        exe2 = os.path.join(OUT, "mjref_hipham")
        subprocess.check_call([a.fc] + FLAGS + ["-module-dir", hmod, "-I", hmod,
                               os.path.join(HERE, "mjref_hipham.f90"), hlib,
                               "-L" + os.path.dirname(hiplib), "-lmatjes_hip",
                               "-Wl,-rpath,$ORIGIN/../../matjes_amd", "-o", exe2])
        print("built", exe2)


if __name__ == "__main__":
    main()
