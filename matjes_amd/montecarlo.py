"""Metropolis Monte Carlo over a temperature list, the loop of montecarlo_run
(src/Metropolis-MC/MonteCarlo.f90:32-172) on the device sweeps of
llg.Hamiltonian.mc_sweeps.

  temperatures   ini_temp (Metropolis-MC/initialize_temperature.f90:12-47):
                 linear from kTini to kTfin, or geometric
  per temperature
    relaxation   N_relaxation x T_relax sweeps (Relaxation.f90:29-79; one sweep
                 = N_spin single-site proposals, the reference's T_relax * N_spin
                 MCsteps)
    measurement  Total_MC_Steps samples, T_auto sweeps before each
                 (MonteCarlo.f90:125-131, measure_add expectation_value.f90:165-205)
    thermo       therm_set (thermo_val.f90:46-91): E, C, M, chi_M and their errors
  output         EM.dat as thermo_print (thermo_val.f90:101-140), the topological
                 and fluctuation columns zero (not computed here)

The Monte Carlo state (spins, cone angle, acceptance counters) carries over
from one temperature to the next as the reference's lattice and track_val do.

parallel_tempering() is the replica-exchange form (src/paratemp-MC/): all
temperatures at once, one replica each, exchanges between neighbouring rungs.
"""
import os

import numpy as np

from . import _lib
from .emio import fortran_e
from .lattice import K_B

SAMPLERS = {"sphere": _lib.MJ_MC_SPHERE, "ising": _lib.MJ_MC_ISING}

EM_HEADER = ['# 1:T', '2:E_av', '3:E_err', '4:C', '5:M', '6:Mx', '7:My', '8:Mz', '9:M_err_x', '10:M_err_y',
             '11:M_err_z', '12:chi_x', '13:chi_y', '14:chi_z', '15:vx', '16:vy', '17:vz', '18:qeuler', '19:Chi_q',
             '20:Q+', '21:Chi_qp', '22:Q-', '23:Chi_qm', '24:l_x', '25:l_y', '26:l_z', '27:Chi_QpQm',
             '28: Re(Mj+Mi-)', '29: Im(Mj+Mi-)', '30: Re(Mi+Mi+)', '31: Im(Mi+Mi+)', '32: Re(Mi+Mi-)',
             '33: Re(Mi+Mj+)', '34: Im(Mi+Mj+)']


def temperatures(kT_ini, kT_fin, n, geometric=False):
    """ini_temp: n values of kT from kT_ini to kT_fin."""
    if n <= 1:
        return np.array([kT_ini], dtype=np.float64)
    i = np.arange(n, dtype=np.float64)
    if geometric:
        return (kT_fin / kT_ini) ** (i / (n - 1)) * kT_ini
    return (kT_fin - kT_ini) / (n - 1) * i + kT_ini


def therm_set(E, M, kT, n_sites):
    """therm_set from the samples of the total energy E (k,) and of the summed
    spins M (k,3): energy, specific heat, magnetisation, susceptibility per site."""
    E = np.asarray(E, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64).reshape(-1, 3)
    n_add = len(E)
    av_site, av_n, av_kt = 1.0 / n_sites, 1.0 / n_add, 1.0 / kT
    e_sum, e_sq = E.sum(), np.sum(E ** 2)
    m_sum, m_sq = M.sum(axis=0), np.sum(M ** 2, axis=0)
    out = {
        "kT": kT,
        "C": (e_sq * av_n - (e_sum * av_n) ** 2) * av_kt ** 2 * av_site,
        "chi_M": (m_sq * av_n - (m_sum * av_n) ** 2) * av_kt * av_site,
        "E": e_sum * av_n * av_site,
        "M": m_sum * av_n * av_site,
        "E_err": 0.0,
        "M_err": np.zeros(3),
        "N_add": n_add,
    }
    if n_add > 1:
        out["E_err"] = np.sqrt(abs(e_sq - e_sum ** 2 * av_n) / (n_add - 1)) * av_site
        out["M_err"] = np.sqrt(np.abs(m_sq - m_sum ** 2 * av_n) / (n_add - 1)) * av_site
    return out


def em_row(t):
    vals = [t["kT"] / K_B, t["E"], t["E_err"], t["C"], float(np.sqrt(np.sum(t["M"] ** 2))), *t["M"], *t["M_err"],
            *t["chi_M"]] + [0.0] * 20
    return "".join(fortran_e(v, 20, 10, 3) + ("  " if k < len(vals) - 1 else "") for k, v in enumerate(vals))


def write_em_dat(path, thermo):
    with open(path, "w") as f:
        f.write("".join("%s%s" % (h, " " * 15) for h in EM_HEADER).rstrip() + "\n")
        for t in thermo:
            f.write(em_row(t) + "\n")


def montecarlo(h, kT_list, n_relaxation=100, T_relax=1, total_mc_steps=100, T_auto=1, sampler="sphere",
               cone=None, seed=0, outdir=None):
    """Run the Monte Carlo loop of montecarlo_run on device Hamiltonian h (its
    current state is the start).  Returns the list of therm_set dicts; writes
    EM.dat into outdir when given."""
    smp = SAMPLERS[sampler] if isinstance(sampler, str) else int(sampler)
    if cone is not None:
        h.mc_set_cone(cone)
    n = h.ncell
    thermo = []
    sweep = 0
    for kT in kT_list:
        n_relax = n_relaxation * T_relax
        if n_relax > 0:
            h.mc_sweeps(n_relax, smp, kT, seed=seed, sweep0=sweep)
            sweep += n_relax
        nmeas = total_mc_steps * T_auto
        E, M = h.mc_sweeps(nmeas, smp, kT, seed=seed, sweep0=sweep, measure_every=T_auto)
        sweep += nmeas
        t = therm_set(E, M, kT, n)
        t["rate"] = h.mc_stats()["rate"]
        thermo.append(t)
    if outdir is not None:
        write_em_dat(os.path.join(outdir, "EM.dat"), thermo)
    return thermo


def parallel_tempering(h, kT_list, n_swap=100, autocor_sweeps=1, relax_sweeps=0, sampler="sphere", cone=1.0, seed=0,
                       outdir=None):
    """The loop of parallel_tempering_run (src/paratemp-MC/parallel_tempering.f90:
    163-181) for one temperature set: replicas of the current state, one per
    temperature, swept in the same launches; every autocor_sweeps sweeps a sample
    of each rung (measure_add) and one exchange of neighbouring rungs
    (reorder_temperature), n_swap times, after relax_sweeps sweeps of relaxation
    with exchanges.  Returns the therm_set dicts by rung (plus swap acceptance);
    writes EM.dat into outdir when given."""
    smp = SAMPLERS[sampler] if isinstance(sampler, str) else int(sampler)
    h.mc_pt_setup(kT_list, cone=cone)
    sweep = 0
    if relax_sweeps > 0:
        h.mc_pt_sweeps(relax_sweeps, smp, seed=seed, sweep0=sweep, exchange_every=autocor_sweeps)
        sweep += relax_sweeps
    E, M = h.mc_pt_sweeps(n_swap * autocor_sweeps, smp, seed=seed, sweep0=sweep, exchange_every=autocor_sweeps,
                          measure_every=autocor_sweeps)
    st = h.mc_pt_state()
    thermo = []
    for q, kT in enumerate(kT_list):
        t = therm_set(E[:, q], M[:, q], kT, h.ncell)
        if q < len(kT_list) - 1:
            t["swap_rate"] = st["swaps_accepted"][q] / max(st["swaps_attempted"][q], 1.0)
        thermo.append(t)
    if outdir is not None:
        write_em_dat(os.path.join(outdir, "EM.dat"), thermo)
    return thermo
