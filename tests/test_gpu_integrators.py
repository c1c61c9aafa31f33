"""The device integrators against closed forms and equilibrium statistics
(MI355X, through the C-ABI).

Independent spins in a static field B along z:
- at T = 0 every spin follows damped precession in closed form, so SIB and
  Depondt are checked against it directly (second order in dt);
- at finite T the in-kernel thermal field (Philox normals, amplitude formed on
  the device) gives <m_z> = L(c x), x = mu_s mu_B B / kT, with c = 1 for SIB
  and Depondt (fluctuation-dissipation amplitude) and c = 2 / 4 for the
  reference's integration 1 / 2 (sigma^2 = alpha kT hbar / dt_stage in the
  last stage only, spindynamics.f90:266-269).
"""
import numpy as np
import pytest

from matjes_amd import lattice as L

pytestmark = pytest.mark.gpu


def _paramagnet_case(n, B, m0):
    lat = L.Lattice(np.eye(3), (1.0, 1.0, 1.0), (n, n, 1), (True, True, False), [("Fe", (0, 0, 0), 1.0)])
    H = L.Hamiltonian(lat)
    H.add_zeeman(B)
    return {"lat": lat, "H": H, "m0": m0}


def _langevin(x):
    return 1.0 / np.tanh(x) - 1.0 / x


@pytest.mark.parametrize("alpha", [0.1, 0.5])
@pytest.mark.parametrize("integ", [5, 6])
def test_damped_precession_closed_form_on_device(integ, alpha):
    from matjes_amd import llg
    n, Bz, t_total = 64, 10.0, 4000.0
    rng = np.random.default_rng(3)
    th0 = rng.uniform(0.2, 2.9, n * n)
    ph0 = rng.uniform(0.0, 2 * np.pi, n * n)
    m0 = np.stack([np.sin(th0) * np.cos(ph0), np.sin(th0) * np.sin(ph0), np.cos(th0)], axis=1)
    w = L.MU_B * Bz / L.HBAR
    th = 2.0 * np.arctan(np.tan(th0 / 2.0) * np.exp(-alpha * w * t_total / (1.0 + alpha * alpha)))
    ph = ph0 + w * t_total / (1.0 + alpha * alpha)
    want = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    err = []
    for dt in (8.0, 4.0):
        with llg.from_case(_paramagnet_case(n, [0.0, 0.0, Bz], m0)) as h:
            h.steps(int(round(t_total / dt)), integ, dt, alpha)
            err.append(np.max(np.abs(h.get_state() - want)))
    assert err[1] < 1e-4, err
    assert 3.5 < err[0] / err[1] < 4.5, err


@pytest.mark.parametrize("integ,c", [(1, 2.0), (2, 4.0), (5, 1.0), (6, 1.0)])
def test_paramagnet_equilibrium_on_device(integ, c):
    """In-kernel thermal field: x = 0.168 (B = 50 T, T = 200 K), 256x256 spins;
    L(x) = 0.056, L(2x) = 0.111, L(4x) = 0.217: a sqrt 2 in sigma moves <m_z>
    by more than 0.04."""
    from matjes_amd import llg
    from oracle import oracle as orc
    n, Bz, T, alpha, dt = 256, 50.0, 200.0, 1.0, 1.0
    kT = T * orc.K_B
    x = L.MU_B * Bz / kT
    with llg.from_case(_paramagnet_case(n, [0.0, 0.0, Bz], np.tile([0.0, 0.0, 1.0], (n * n, 1)))) as h:
        h.steps(2000, integ, dt, alpha, kT=kT, seed=11, step0=0)
        acc = []
        for k in range(20):
            h.steps(200, integ, dt, alpha, kT=kT, seed=11, step0=2000 + 200 * k)
            acc.append(h.get_state()[:, 2].mean())
    assert abs(np.mean(acc) - _langevin(c * x)) < 0.01, (np.mean(acc), _langevin(c * x))
