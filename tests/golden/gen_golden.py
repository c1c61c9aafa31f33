#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference build.

Runs oracle/_ref/mjref_driver (the reference's own Fortran, built by
oracle/ref/build_ref.py) on small inputs in the reference's `input` format and
stores inputs + outputs as .npz (plain arrays, no pickles):

  setup    JSON string: lattice/Hamiltonian/dynamics parameters (what
           matjes_amd.workloads.build_case needs to rebuild the same terms)
  m0       (N,3) initial state (magnetic.inp)
  bmap     (N,3) per-site B order parameter (Bfield.inp), when used
  field    (3N,) reference B_eff of m0       [field cases]
  energy   ()    reference total energy of m0 [field cases]
  term_names, term_energies  energy of each term of the resolved Hamiltonian [field cases]
  steps    (K,)  output step numbers          [traj cases]
  E        (K,)  E/Ncell after those steps    [traj cases]
  m        (K,N,3) states after those steps   [traj cases]
  noise_sigma  the thermal-field sigma the reference computed (sigma.out of
           the driver) when T>0; the standard normals it scaled are the
           oracle's Philox/Box-Muller draws for setup["noise_seed"], regenerated
           by the tests (oracle.reference_normals)

This script runs only in the build container (it needs /root/reference for the
build); the fixtures it writes are what travels.
Usage: python tests/golden/gen_golden.py [case ...]
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from matjes_amd import workloads  # noqa: E402
from oracle import oracle as orc  # noqa: E402

DRIVER = os.path.join(REPO, "oracle", "_ref", "mjref_driver")


def input_text(setup, mode):
    s = setup
    geo = s["geometry"]
    lines = ["simulation magnet-dynamics", "Hamiltonian_mode 5"]
    per = " ".join("T" if p else "F" for p in s["periodic"])
    lines.append("Periodic_log " + per)
    lines.append("Nsize %d %d %d" % tuple(s["size"]))
    a = s.get("alat", 1.0)
    lines.append("alat %r %r %r" % (a, a, a))
    lines.append("lattice")
    if geo == "triangular":
        lines += ["1.0 0.0 0.0", "0.5 0.86602540378443864676 0.0", "0.0 0.0 1.0"]
    else:
        lines += ["1.0 0.0 0.0", "0.0 1.0 0.0", "0.0 0.0 1.0"]
    has_med = s.get("D", 0.0) != 0.0
    lines.append("")
    lines.append("atomtypes %d" % (2 if has_med else 1))
    lines.append("Fe %r 0.0 0.0 F" % s["moment"])
    if has_med:
        lines.append("Ir 0.0 0.0 0.0 F")
    lines.append("")
    lines.append("atoms %d" % (2 if has_med else 1))
    lines.append("Fe 0.0 0.0 0.0")
    if has_med:
        lines.append("Ir 0.0 0.0 %r" % s.get("mediator_dz", -0.5))
    lines.append("")
    J = s.get("J", [])
    if J:
        lines.append("magnetic_J")
        for i, j in enumerate(J):
            lines.append("1 1 %d %r" % (i + 1, j))
        lines.append("")
    if has_med:
        lines.append("magnetic_D")
        lines.append("1 1 2 1 %r" % s["D"])
        lines.append("")
    if s.get("K", 0.0) != 0.0:
        e = s.get("K_axis", [0.0, 0.0, 1.0])
        lines.append("magnetic_anisotropy")
        lines.append("1 %r %r %r %r" % (e[0], e[1], e[2], s["K"]))
        lines.append("")
    B = s.get("B", [0.0, 0.0, 0.0])
    lines.append("H_ext %r %r %r" % tuple(B))
    if s.get("bmap_sigma"):
        lines.append("enable_H T")
    if s.get("dipolar"):
        lines.append("mag_dip_use T")
        lines.append("mag_dip_fft F")
        lines.append("mag_dip_period_cut %d %d %d" % tuple(s.get("dip_cut", [1, 1, 1])))
    lines.append("Tini %r" % s.get("T", 0.0))
    if mode == "traj":
        lines.append("integration %d" % s["integration"])
        if s["integration"] == 4:
            lines.append("N_order %d" % s.get("N_order", 4))
        lines.append("timestep %r" % s["dt"])
        lines.append("duration %d" % s["nsteps"])
        lines.append("damping %r" % s["alpha"])
        lines.append("Efreq %d" % s["efreq"])
        if s.get("discretization"):
            lines.append("update_time T")
            lines.append("discretization %r" % s["discretization"])
    return "\n".join(lines) + "\n"


def run_case(name, setup, mode):
    case = workloads.build_case(setup)
    m0 = case["m0"]
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "input"), "w").write(input_text(setup, mode))
        np.savetxt(os.path.join(d, "magnetic.inp"), m0, fmt="%.17e")
        if case.get("bmap") is not None:
            np.savetxt(os.path.join(d, "Bfield.inp"), case["bmap"], fmt="%.17e")
        out = {"setup": np.array(json.dumps(setup)), "m0": m0}
        if case.get("bmap") is not None:
            out["bmap"] = case["bmap"]
        noisy = mode == "traj" and setup.get("T", 0.0) > 0.0
        if noisy:
            # standard normals only: the reference build scales them by the sigma it
            # computes itself (spindynamics.f90:269) and writes that sigma to sigma.out
            with open(os.path.join(d, "normals.bin"), "wb") as fh:
                for j in range(setup["nsteps"]):
                    fh.write(orc.reference_normals(setup, case["lat"].ncell, j).astype("<f8").tobytes())
        r = subprocess.run(["bash", "-c", "ulimit -s unlimited && exec %s %s" % (DRIVER, mode)],
                           cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference run failed for %s:\n%s\n%s" % (name, r.stdout[-3000:], r.stderr[-3000:]))
        if mode == "field":
            v = np.loadtxt(os.path.join(d, "field.out"))
            out["energy"] = np.array(v[0])
            out["field"] = v[1:]
            lines = open(os.path.join(d, "eres.out")).read().splitlines()
            out["term_names"] = np.array([t.strip() for t in lines[0::2]])
            out["term_energies"] = np.array([float(t) for t in lines[1::2]])
        else:
            v = np.loadtxt(os.path.join(d, "traj.out"))
            n3 = 3 * case["lat"].ncell
            rec = v.reshape(-1, n3 + 2)
            out["steps"] = rec[:, 0].astype(np.int64)
            out["E"] = rec[:, 1]
            out["m"] = rec[:, 2:].reshape(len(rec), -1, 3)
            if noisy:
                sig = np.atleast_1d(np.loadtxt(os.path.join(d, "sigma.out")))
                assert len(sig) == setup["nsteps"] and np.all(sig == sig[0])
                out["noise_sigma"] = np.array(sig[0])
            if setup.get("discretization"):
                out["timesteps"] = np.atleast_1d(np.loadtxt(os.path.join(d, "ts.out")))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name)


CASES = {
    # square lattice, Hamiltonian of the 64x64 deterministic workload (J, interfacial D, K, B)
    "field_sq8": ("field", dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=11)),
    # triangular lattice, Hamiltonian of the 1024x1024 stochastic workload (J1-3, D, K, B)
    "field_tri8": ("field", dict(workloads.TRIANGULAR_J3DKB, size=[8, 8, 1], seed=12)),
    # per-site field map entering through the Zeeman order parameter
    "field_sq8_bmap": ("field", dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=13, bmap_sigma=0.5)),
    # dipolar (dense reference) on open and on in-plane periodic films
    "field_dip_open": ("field", dict(workloads.FILM_DIP, size=[4, 3, 2], periodic=[False, False, False], seed=14)),
    "field_dip_ttf": ("field", dict(workloads.FILM_DIP, size=[4, 4, 2], periodic=[True, True, False], seed=15)),
    # trajectories with the reference's own Euler (1) and two-half-step Heun (2)
    "traj_sq8_euler": ("traj", dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=21, integration=1,
                                    dt=1.0, alpha=0.1, nsteps=300, efreq=100)),
    "traj_sq8_heun": ("traj", dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=22, integration=2,
                                   dt=1.0, alpha=0.1, nsteps=300, efreq=100)),
    "traj_tri8_heun_T10": ("traj", dict(workloads.TRIANGULAR_J3DKB, size=[8, 8, 1], seed=23, integration=2,
                                        dt=0.5, alpha=0.05, nsteps=200, efreq=50, T=10.0, noise_seed=7)),
    "traj_tri8_euler_T10": ("traj", dict(workloads.TRIANGULAR_J3DKB, size=[8, 8, 1], seed=27, integration=1,
                                         dt=0.5, alpha=0.05, nsteps=200, efreq=50, T=10.0, noise_seed=8)),
    "traj_sq8_heun_bmap": ("traj", dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=24, integration=2,
                                        dt=1.0, alpha=0.2, nsteps=200, efreq=100, bmap_sigma=0.5)),
    # the reference's Runge-Kutta order N (integration 4, N_order 4, 3 and 7; T = 0 only)
    "traj_sq8_rk4": ("traj", dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=41, integration=4, N_order=4,
                                  dt=1.0, alpha=0.1, nsteps=300, efreq=100)),
    "traj_tri8_rk3": ("traj", dict(workloads.TRIANGULAR_J3DKB, size=[8, 8, 1], seed=42, integration=4, N_order=3,
                                   dt=0.5, alpha=0.05, nsteps=200, efreq=100)),
    "traj_sq8_rk7": ("traj", dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=43, integration=4, N_order=7,
                                  dt=1.0, alpha=0.1, nsteps=200, efreq=100, bmap_sigma=0.5)),
    # 10,000 steps at T=0 on 32x32 (1e-10 trajectory parity)
    "traj_sq32_heun_10k": ("traj", dict(workloads.SQUARE_JDKB, size=[32, 32, 1], seed=31, integration=2,
                                        dt=1.0, alpha=0.1, nsteps=10000, efreq=2500)),
    "traj_sq32_euler_10k": ("traj", dict(workloads.SQUARE_JDKB, size=[32, 32, 1], seed=32, integration=1,
                                         dt=1.0, alpha=0.1, nsteps=10000, efreq=2500)),
    # 10,000 steps at T = 10 K on 32x32 with the Hamiltonian and integrator of the
    # triangular 1024x1024 workload (normals regenerated from noise_seed)
    "traj_tri32_heun_T10_10k": ("traj", dict(workloads.TRIANGULAR_J3DKB, size=[32, 32, 1], seed=33, integration=2,
                                             dt=0.5, alpha=0.05, nsteps=10000, efreq=2500, T=10.0, noise_seed=9)),
    # adaptive timestep (update_time T, discretization 15) with Heun and Euler
    "traj_sq8_heun_update_time": ("traj", dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=25, integration=2,
                                               dt=1.0, alpha=0.1, nsteps=150, efreq=50, discretization=15.0)),
    "traj_sq8_euler_update_time": ("traj", dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=26, integration=1,
                                                dt=1.0, alpha=0.1, nsteps=150, efreq=50, discretization=15.0)),
}


def main(argv):
    if not os.path.isfile(DRIVER):
        sys.exit("reference driver missing: run python oracle/ref/build_ref.py first")
    names = argv or [k for k in CASES]
    for n in names:
        mode, setup = CASES[n]
        run_case(n, setup, mode)


if __name__ == "__main__":
    main(sys.argv[1:])
