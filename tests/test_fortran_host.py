"""Fortran host code (iso_c_binding module + driver) over the C-ABI."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO

EXE = os.path.join(REPO, "matjes_amd", "fortran", "mjf_llg")


def _write_case(path, c, integ, nsteps, dt, alpha, kT, seed):
    H = c["H"]
    lat = c["lat"]
    r, cc, v = H.coo_mm()
    zr, zc, zv = H.coo_zeeman()
    with open(path, "wb") as f:
        np.array(list(lat.dim_lat) + [int(p) for p in lat.periodic] + [integ, nsteps], np.int32).tofile(f)
        np.array([dt, alpha, kT], np.float64).tofile(f)
        np.array([seed], np.int64).tofile(f)
        np.array([len(v)], np.int64).tofile(f)
        r.astype(np.int32).tofile(f)
        cc.astype(np.int32).tofile(f)
        v.astype(np.float64).tofile(f)
        np.array([len(zv)], np.int64).tofile(f)
        zr.astype(np.int32).tofile(f)
        zc.astype(np.int32).tofile(f)
        zv.astype(np.float64).tofile(f)
        H.b_field_sites().reshape(-1).astype(np.float64).tofile(f)
        c["m0"].reshape(-1).astype(np.float64).tofile(f)


def test_fortran_binding_compiles(tmp_path):
    from matjes_amd import build
    fc = build.flang()
    if fc is None:
        pytest.skip("amdflang not present")
    mod = os.path.join(REPO, "matjes_amd", "fortran", "m_matjes_hip.f90")
    subprocess.check_call([fc, "-c", mod, "-module-dir", str(tmp_path), "-o", str(tmp_path / "m.o")])


def test_hip_backend_compiles_against_the_reference_modules(tmp_path):
    """The t_H_hip backend type extends the reference's t_H_coo_based and
    fft_H_hip its fft_H: both compile against the modules of the reference
    build (where that build exists)."""
    from matjes_amd import build
    fc = build.flang()
    refmod = os.path.join(REPO, "oracle", "_ref", "mod")
    if fc is None or not os.path.isfile(os.path.join(refmod, "m_h_coo_based.mod")):
        pytest.skip("amdflang or the reference build not present")
    fdir = os.path.join(REPO, "matjes_amd", "fortran")
    flags = ["-cpp", "-O2", "-fopenmp", "-DCPP_SCRIPT", "-module-dir", str(tmp_path), "-I", str(tmp_path), "-I", refmod]
    subprocess.check_call([fc] + flags + ["-c", os.path.join(fdir, "m_matjes_hip.f90"), "-o", str(tmp_path / "m.o")])
    subprocess.check_call([fc] + flags + ["-c", os.path.join(fdir, "ham_hip.f90"), "-o", str(tmp_path / "h.o")])
    subprocess.check_call([fc] + flags + ["-c", os.path.join(fdir, "fft_ham_hip.f90"), "-o", str(tmp_path / "f.o")])
    assert (tmp_path / "m_h_hip.mod").exists() and (tmp_path / "m_fft_h_hip.mod").exists()


@pytest.mark.gpu
def test_fortran_driver_heun_vs_oracle(oracle_lib, tmp_path):
    from matjes_amd import workloads
    if not os.path.isfile(EXE):
        pytest.fail("Fortran driver not built (matjes_amd/build.py build_fortran)")
    c = workloads.build_case(dict(workloads.TRIANGULAR_J3DKB, size=[16, 16, 1], seed=21))
    nsteps = 60
    _write_case(str(tmp_path / "case.bin"), c, 2, nsteps, 0.5, 0.05, 0.0, 0)
    r = subprocess.run([EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(str(tmp_path / "out.bin"), np.float64)
    n3 = 3 * c["lat"].ncell
    B, E0, E, m = out[:n3], out[n3], out[n3 + 1:n3 + 1 + nsteps], out[n3 + 1 + nsteps:]
    o = oracle_lib.Oracle(c["H"])
    Bo = o.field(c["m0"]).reshape(-1)
    assert np.max(np.abs(B - Bo)) <= 1e-14 * np.max(np.abs(Bo))
    assert abs(E0 - o.energy(c["m0"])) <= 1e-13 * abs(E0)
    mo, Eo = o.steps(c["m0"], nsteps, 2, 0.5, 0.05, energies=True)
    assert np.max(np.abs(m.reshape(-1, 3) - mo)) <= 1e-12
    assert np.max(np.abs(E - Eo)) <= 1e-12 * max(np.max(np.abs(Eo)), 1e-3)
