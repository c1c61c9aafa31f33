"""Pins of the oracle's SIB and Depondt integrators and of every integrator's
thermal amplitude, independent of the oracle's own formulas (CPU).

- A single spin in a static field B along z follows damped precession in
  closed form for the reference's LLG (propagator.f90:13-40, energies in eV):
      tan(theta/2) = tan(theta0/2) exp(-alpha w t / (1 + alpha^2)),
      phi          = phi0 + w t / (1 + alpha^2),       w = mu_s mu_B B / hbar.
  SIB and Depondt converge to it at second order.
- Non-interacting spins in B at temperature T equilibrate to the Langevin
  function <m.B^> = L(c x), x = mu_s mu_B B / kT.  SIB and Depondt use the
  fluctuation-dissipation amplitude 2 alpha kT hbar / dt (c = 1).  The
  reference's integration 1 and 2 draw sigma^2 = alpha kT hbar / dt_stage
  (spindynamics.f90:266-269) and apply it in the last stage only, so a step of
  length dt receives 1/2 (Euler, dt_stage = dt) or 1/4 (two half-steps,
  dt_stage = dt/2) of the fluctuation-dissipation impulse: c = 2 and c = 4.
- Mistakes planted in a copy of the oracle's source (a dropped 1/(1+alpha^2),
  a doubled rotation angle, a thermal amplitude off by sqrt 2) fail these
  checks, so the checks have the power to catch them.
"""
import os
import subprocess

import numpy as np
import pytest

from matjes_amd import lattice as L

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "mjoracle.c")


def _single_spin(B, moment=1.0):
    lat = L.Lattice(np.eye(3), (1.0, 1.0, 1.0), (1, 1, 1), (False, False, False), [("Fe", (0, 0, 0), moment)])
    H = L.Hamiltonian(lat)
    H.add_zeeman(B)
    return H


def _paramagnet(n, B, moment=1.0):
    lat = L.Lattice(np.eye(3), (1.0, 1.0, 1.0), (n, n, 1), (True, True, False), [("Fe", (0, 0, 0), moment)])
    H = L.Hamiltonian(lat)
    H.add_zeeman(B)
    return H


def _langevin(x):
    return 1.0 / np.tanh(x) - 1.0 / x


B_Z, THETA0, PHI0, T_TOTAL = 10.0, 1.2, 0.3, 4000.0


def _closed_form(alpha, t):
    w = L.MU_B * B_Z / L.HBAR
    th = 2.0 * np.arctan(np.tan(THETA0 / 2.0) * np.exp(-alpha * w * t / (1.0 + alpha * alpha)))
    ph = PHI0 + w * t / (1.0 + alpha * alpha)
    return np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])


def _precession_errors(orc, integ, alpha, dts=(8.0, 4.0)):
    o = orc.Oracle(_single_spin([0.0, 0.0, B_Z]))
    m0 = np.array([[np.sin(THETA0) * np.cos(PHI0), np.sin(THETA0) * np.sin(PHI0), np.cos(THETA0)]])
    want = _closed_form(alpha, T_TOTAL)
    return [np.max(np.abs(o.steps(m0, int(round(T_TOTAL / dt)), integ, dt, alpha)[0] - want)) for dt in dts]


@pytest.mark.parametrize("alpha", [0.1, 0.5])
@pytest.mark.parametrize("integ", [5, 6])
def test_damped_precession_closed_form(oracle_lib, integ, alpha):
    """SIB (5) and Depondt (6) against the closed form: error below 1e-4 at
    dt = 4 fs (w dt = 3.5e-3) and second order (ratio 4 on halving dt)."""
    e8, e4 = _precession_errors(oracle_lib, integ, alpha)
    assert e4 < 1e-4, e4
    assert 3.5 < e8 / e4 < 4.5, (e8, e4)


def _equilibrium(orc, integ, T=200.0, B=50.0, alpha=1.0, dt=1.0, n=64):
    """<m_z> of n x n independent spins (B along z) after 2000 fs, mean of 20
    samples 200 fs apart; returns (mean, standard error, x)."""
    kT = T * orc.K_B
    o = orc.Oracle(_paramagnet(n, [0.0, 0.0, B]))
    m = np.tile([0.0, 0.0, 1.0], (n * n, 1))
    nw, ns = int(round(2000.0 / dt)), int(round(200.0 / dt))
    m = o.steps(m, nw, integ, dt, alpha, kT=kT, seed=1, step0=0)
    acc = []
    for k in range(20):
        m = o.steps(m, ns, integ, dt, alpha, kT=kT, seed=1, step0=nw + ns * k)
        acc.append(m[:, 2].mean())
    acc = np.array(acc)
    return acc.mean(), acc.std() / np.sqrt(len(acc)), L.MU_B * B / kT


# c of <m.B^> = L(c x): the fraction of the fluctuation-dissipation impulse each scheme applies, inverted
EQUIL_C = {1: 2.0, 2: 4.0, 5: 1.0, 6: 1.0}


@pytest.mark.parametrize("integ", [1, 2, 5, 6])
def test_paramagnet_equilibrium(oracle_lib, integ):
    """Thermal amplitudes: x = 0.168 gives L(x) = 0.056, L(2x) = 0.111,
    L(4x) = 0.217, L(8x) = 0.39; the sampling error is about 0.002."""
    mz, err, x = _equilibrium(oracle_lib, integ)
    assert abs(mz - _langevin(EQUIL_C[integ] * x)) < 0.01, (mz, err, _langevin(EQUIL_C[integ] * x))


@pytest.fixture
def planted(monkeypatch, tmp_path):
    """Load the oracle compiled from a copy of its source with one expression replaced."""
    from oracle import oracle as orc

    def make(old, new):
        src = open(ORACLE_SRC).read()
        assert src.count(old) == 1, old
        c = tmp_path / "mjoracle_planted.c"
        c.write_text(src.replace(old, new))
        so = tmp_path / "liboracle_planted.so"
        subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
                               str(c), "-o", str(so), "-lm"])
        monkeypatch.setattr(orc, "LIB", str(so))
        monkeypatch.setattr(orc, "_lib", None)
        return orc
    return make


PLANTS = {
    "sib_drops_damping_denominator": (5, "double f = -0.5 * dt / ((1.0 + alpha * alpha) * OR_HBAR);",
                                      "double f = -0.5 * dt / OR_HBAR;"),
    "depondt_drops_damping_denominator": (6, "bt[k] = (B[k] + alpha * c[k]) / den;", "bt[k] = (B[k] + alpha * c[k]);"),
    "depondt_doubles_rotation_angle": (6, "double th = bn * dt / OR_HBAR;", "double th = 2.0 * bn * dt / OR_HBAR;"),
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_integrator_mistake_fails_closed_form(planted, plant):
    integ, old, new = PLANTS[plant]
    orc = planted(old, new)
    for alpha in (0.1, 0.5):
        e8, e4 = _precession_errors(orc, integ, alpha)
        if e4 >= 1e-4 or not 3.5 < e8 / e4 < 4.5:
            return
    pytest.fail("planted mistake %s passed the closed-form check" % plant)


SIGMA_PLANTS = {
    # the full timestep instead of timestep / N_loop in the reference scheme's sigma
    "heun_sigma_from_full_timestep": (2, "sqrt(d->damping * d->kT * OR_HBAR / dt), bt);",
                                      "sqrt(d->damping * d->kT * OR_HBAR / ts), bt);"),
    # the reference's amplitude (no factor 2) in the SIB/Depondt scheme
    "sib_sigma_without_factor_two": (5, "sqrt(2.0 * d->damping * d->kT * OR_HBAR / dt), bt);",
                                     "sqrt(d->damping * d->kT * OR_HBAR / dt), bt);"),
}


@pytest.mark.parametrize("plant", sorted(SIGMA_PLANTS))
def test_planted_thermal_amplitude_mistake_fails_equilibrium(planted, plant):
    integ, old, new = SIGMA_PLANTS[plant]
    orc = planted(old, new)
    mz, err, x = _equilibrium(orc, integ)
    assert abs(mz - _langevin(EQUIL_C[integ] * x)) >= 0.01, (mz, _langevin(EQUIL_C[integ] * x))
