"""Device Metropolis Monte Carlo (through the C-ABI) against the oracle, closed
forms and the reference's own Monte Carlo test output."""
import os

import numpy as np
import pytest
from scipy.special import ellipk

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

SPHERE, ISING = 0, 1


def _lattice_case(nx, ny, periodic, J=0.0, K=0.0, B=None, moment=1.0, spins=(0.0, 0.0, 1.0)):
    from matjes_amd import lattice as L
    lat = L.square_film(nx, ny, 1, alat=1.0, moment=moment, mediator_dz=None, periodic=periodic)
    H = L.Hamiltonian(lat)
    if J:
        H.add_exchange(1, J)
    if K:
        H.add_anisotropy([0.0, 0.0, 1.0], K)
    if B is not None:
        H.add_zeeman(np.asarray(B, dtype=np.float64))
    m0 = np.tile(np.asarray(spins, dtype=np.float64), (lat.ncell, 1))
    return {"lat": lat, "H": H, "m0": m0}


def _mean_err(x, nb=20):
    x = np.asarray(x)
    b = x[:len(x) // nb * nb].reshape(nb, -1).mean(axis=1)
    return x.mean(), b.std(ddof=1) / np.sqrt(nb)


def _dev(case, form="stencil"):
    from matjes_amd import llg
    return llg.from_case(case, form=form)


CASES = [("TRIANGULAR_J3DKB", (32, 32, 1), (True, True, False)),
         ("SQUARE_JDKB_BMAP", (24, 20, 1), (True, True, False)),
         ("TRIANGULAR_J3DKB", (21, 15, 1), (False, False, False)),
         ("FILM", (16, 16, 4), (True, True, False))]


def _case(setup, size, periodic, seed=3):
    from matjes_amd import workloads
    if setup == "SQUARE_JDKB_BMAP":
        s = dict(workloads.SQUARE_JDKB, bmap_sigma=0.5)
    elif setup == "FILM":
        s = dict(workloads.FILM_DIP, dipolar=False, B=[0.0, 0.0, 1.0])
    else:
        s = getattr(workloads, setup)
    return workloads.build_case(dict(s, size=list(size), periodic=list(periodic), seed=seed))


@pytest.mark.parametrize("setup,size,periodic", CASES)
def test_energy_single_vs_oracle(oracle_lib, setup, size, periodic):
    c = _case(setup, size, periodic)
    o = oracle_lib.Oracle(c["H"])
    rng = np.random.default_rng(2)
    n = c["lat"].ncell
    sites = rng.integers(1, n + 1, size=64)
    new = rng.normal(size=(64, 3))
    new /= np.linalg.norm(new, axis=1)[:, None]
    with _dev(c) as h:
        e_cur = h.energy_single(sites)
        e_new = h.energy_single(sites, new)
        e0 = h.energy()
        col = h.mc_stats()["colouring"]
    ref_cur = o.energy_single(c["m0"], sites)
    ref_new = o.energy_single(c["m0"], sites, new)
    scale = np.max(np.abs(ref_cur))
    assert np.max(np.abs(e_cur - ref_cur)) <= 1e-13 * scale
    assert np.max(np.abs(e_new - ref_new)) <= 1e-13 * scale
    assert col == o.mc_colouring()
    # the energy change of one site equals the change of the device's total energy
    with _dev(c) as h:
        for k in range(4):
            m2 = c["m0"].copy()
            m2[sites[k] - 1] = new[k]
            h.set_state(m2)
            d_tot = h.energy() - e0
            assert abs((e_new[k] - e_cur[k]) - d_tot) <= 1e-11 * max(abs(d_tot), 1e-3 * abs(e0) / n)


@pytest.mark.parametrize("sampler", [SPHERE, ISING])
@pytest.mark.parametrize("setup,size,periodic", CASES)
def test_sweeps_vs_oracle(oracle_lib, setup, size, periodic, sampler):
    """Same Philox draws, same colour classes: the device sweeps and the oracle's
    restatement agree state for state (acceptance decisions identical)."""
    c = _case(setup, size, periodic)
    kT = 10.0 * oracle_lib.K_B
    o = oracle_lib.Oracle(c["H"])
    st = np.array([0.4, np.cos(0.4), 0.0, 0.0, 0.0])
    m_ref, E_ref, M_ref = o.mc_sweeps(c["m0"], 24, sampler, kT, seed=17, sweep0=5, state=st, measure_every=6)
    with _dev(c) as h:
        h.mc_set_cone(0.4)
        E, M = h.mc_sweeps(24, sampler, kT, seed=17, sweep0=5, measure_every=6)
        m = h.get_state()
        stats = h.mc_stats()
    assert stats["accepted"] == st[2] and stats["proposed"] == st[3]
    assert abs(stats["cone"] - st[0]) <= 1e-12
    assert np.max(np.abs(m - m_ref)) <= 1e-12
    assert np.max(np.abs(E - E_ref)) <= 1e-12 * np.max(np.abs(E_ref))
    assert np.max(np.abs(M - M_ref)) <= 1e-9
    assert 0.0 < stats["accepted"] < stats["proposed"]


@pytest.mark.parametrize("frac", [0.8, 1.5])
def test_ising_256_onsager(frac):
    from matjes_amd import llg
    J = 0.5e-3
    Jp = 2.0 * J
    c = _lattice_case(256, 256, (True, True, False), J=J)
    kTc = 2.0 * Jp / np.log(1.0 + np.sqrt(2.0))
    kT = frac * kTc
    K = Jp / kT
    k = 2.0 * np.sinh(2 * K) / np.cosh(2 * K) ** 2
    u = -Jp / np.tanh(2 * K) * (1.0 + 2.0 / np.pi * (2.0 * np.tanh(2 * K) ** 2 - 1.0) * ellipk(k * k))
    with llg.from_case(c, form="stencil") as h:
        h.mc_sweeps(500, ISING, kT, seed=5)
        E, M = h.mc_sweeps(2000, ISING, kT, seed=5, sweep0=500, measure_every=2)
    n = c["lat"].ncell
    e, se = _mean_err(E / n)
    assert abs(e - u) <= 5 * se + 5e-4 * abs(u), (e, se, u)
    if frac < 1.0:
        mag, sm = _mean_err(np.abs(M[:, 2]) / n)
        assert abs(mag - (1.0 - np.sinh(2 * K) ** -4) ** 0.125) <= 5 * sm + 1e-3


def test_ising_against_reference_ptmc_run(oracle_lib):
    """tests/serial/PTMC of the reference: mean energy per site at each of its 40
    temperatures within the run's error bars."""
    from matjes_amd import llg
    ref = np.loadtxt(os.path.join(GOLDEN, "legacy_ptmc_ising20.txt"))
    c = _lattice_case(20, 20, (True, True, False), J=0.5e-3, K=5e-7, moment=3.0)
    n = c["lat"].ncell
    with llg.from_case(c, form="stencil") as h:
        for T, e_ref, e_err, _, _ in ref:
            h.set_state(c["m0"])
            h.mc_sweeps(500, ISING, T * oracle_lib.K_B, seed=int(T * 100))
            E, _ = h.mc_sweeps(4000, ISING, T * oracle_lib.K_B, seed=int(T * 100), sweep0=500, measure_every=4)
            e, se = _mean_err(E / n)
            assert abs(e - e_ref) <= 4.0 * np.hypot(e_err, se) + 1e-9, (T, e, se, e_ref, e_err)


def test_heisenberg_ring_and_paramagnet():
    from matjes_amd import llg
    J = 1.0e-3
    Jp = 2.0 * J
    c = _lattice_case(4096, 1, (True, False, False), J=J)
    with llg.from_case(c, form="stencil") as h:
        for K in (0.5, 2.0):
            h.set_state(c["m0"])
            h.mc_set_cone(0.5)
            h.mc_sweeps(500, SPHERE, Jp / K, seed=3)
            E, _ = h.mc_sweeps(4000, SPHERE, Jp / K, seed=3, sweep0=500, measure_every=4)
            e, se = _mean_err(E / 4096)
            assert abs(e + Jp * (1.0 / np.tanh(K) - 1.0 / K)) <= 5 * se + 5e-4 * Jp, (K, e, se)
    c = _lattice_case(256, 256, (True, True, False), B=[2.0, 0.0, 0.0], moment=2.0, spins=(0.0, 1.0, 0.0))
    mu_b = abs(c["H"].zeeman[0, 0]) * 2.0
    with llg.from_case(c, form="stencil") as h:
        for x in (0.5, 3.0):
            h.set_state(c["m0"])
            h.mc_set_cone(1.0)
            h.mc_sweeps(200, SPHERE, mu_b / x, seed=9)
            _, M = h.mc_sweeps(600, SPHERE, mu_b / x, seed=9, sweep0=200, measure_every=2)
            mx, sx = _mean_err(M[:, 0] / c["lat"].ncell)
            assert abs(mx - (1.0 / np.tanh(x) - 1.0 / x)) <= 5 * sx + 1e-3, (x, mx, sx)


def test_full_size_sweeps_properties():
    """1024x1024 triangular lattice (J1-3, DMI, K, B) at 10 K: unit spins, finite
    energies, a non-trivial acceptance rate, T = 0 never raises the energy."""
    from matjes_amd import llg, workloads
    c = workloads.build_case(dict(workloads.TRIANGULAR_J3DKB, size=[1024, 1024, 1], seed=1))
    with llg.from_case(c, form="stencil") as h:
        h.mc_set_cone(0.5)
        E, M = h.mc_sweeps(8, SPHERE, 10.0 * 8.6173324e-5, seed=1, measure_every=1)
        m = h.get_state()
        st = h.mc_stats()
        E0, _ = h.mc_sweeps(4, SPHERE, 0.0, seed=2, sweep0=8, measure_every=1)
    assert st["colouring"] == (8, 3, 0)
    assert np.max(np.abs(np.linalg.norm(m, axis=1) - 1.0)) <= 1e-13
    assert np.all(np.isfinite(E)) and E[-1] < E[0]
    assert 0.05 < st["rate"] < 0.95
    assert np.all(np.diff(np.concatenate([[E[-1]], E0])) <= 1e-12 * abs(E[-1]))


def test_refusals():
    from matjes_amd import _lib, llg, workloads
    c = workloads.build_case(dict(workloads.FILM_DIP, size=[8, 8, 2], seed=1))
    with llg.from_case(c, form="stencil") as h:
        with pytest.raises(_lib.MatjesError):
            h.mc_sweeps(1, SPHERE, 1e-3)
        with pytest.raises(_lib.MatjesError):
            h.energy_single([1])
    c = workloads.build_case(dict(workloads.SQUARE_JDKB, size=[12, 10, 1], seed=4))
    r, cc, v = c["H"].coo_mm()
    v = v.copy()
    v[::7] *= 1.1
    with llg.Hamiltonian(c["lat"].dim_lat, c["lat"].periodic) as h:
        h.add_coo(1, 1, r, cc, v)
        h.set_state(c["m0"])
        with pytest.raises(_lib.MatjesError):
            h.mc_sweeps(1, ISING, 1e-3)
        with pytest.raises(_lib.MatjesError):
            h.energy_single([0])


def test_montecarlo_driver_ptmc(oracle_lib, tmp_path):
    """montecarlo_run's loop (relaxation, samples every T_auto sweeps, therm_set,
    EM.dat) on the reference's PTMC system at a few of its temperatures."""
    from matjes_amd import llg, montecarlo
    ref = np.loadtxt(os.path.join(GOLDEN, "legacy_ptmc_ising20.txt"))
    rows = ref[[6, 14, 22, 30]]
    c = _lattice_case(20, 20, (True, True, False), J=0.5e-3, K=5e-7, moment=3.0)
    with llg.from_case(c, form="stencil") as h:
        th = montecarlo.montecarlo(h, rows[:, 0] * oracle_lib.K_B, n_relaxation=10, T_relax=50, total_mc_steps=2000,
                                   T_auto=2, sampler="ising", seed=4, outdir=str(tmp_path))
    em = np.loadtxt(tmp_path / "EM.dat")
    assert em.shape == (4, 34)
    assert np.allclose(em[:, 0], rows[:, 0], rtol=1e-9)
    for t, r in zip(th, rows):
        assert abs(t["E"] - r[1]) <= 4.0 * np.hypot(r[2], 3.0 * t["E_err"] / np.sqrt(t["N_add"])) + 1e-9, (t, r)
    assert np.allclose(em[:, 1], [t["E"] for t in th], rtol=1e-9)


def test_pt_one_replica_is_the_plain_sweep():
    from matjes_amd import llg
    c = _case("TRIANGULAR_J3DKB", (32, 32, 1), (True, True, False))
    kT = 10.0 * 8.6173324e-5
    with llg.from_case(c, form="stencil") as h:
        h.mc_set_cone(0.4)
        h.mc_sweeps(10, SPHERE, kT, seed=7, sweep0=3)
        m_plain = h.get_state()
        st_plain = h.mc_stats()
        h.set_state(c["m0"])
        h.mc_pt_setup([kT], cone=0.4)
        h.mc_pt_sweeps(10, SPHERE, seed=7, sweep0=3)
        m_pt = h.mc_pt_get_state(0)
        st = h.mc_pt_state()["replicas"][0]
    assert np.array_equal(m_pt, m_plain)
    assert st[1] == st_plain["accepted"] and st[2] == st_plain["proposed"] and st[0] == st_plain["cone"]


@pytest.mark.parametrize("sampler", [SPHERE, ISING])
@pytest.mark.parametrize("setup,size,periodic", [CASES[0], CASES[1]])
def test_pt_vs_oracle(oracle_lib, setup, size, periodic, sampler):
    """Replicas, exchanges and samples by rung against the oracle's restatement."""
    c = _case(setup, size, periodic)
    ladder = np.geomspace(2.0, 60.0, 6) * oracle_lib.K_B
    o = oracle_lib.Oracle(c["H"])
    pt, E_ref, M_ref = o.mc_pt(c["m0"], ladder, 16, sampler, seed=11, sweep0=2, exchange_every=2, measure_every=4,
                               cone=0.6)
    from matjes_amd import llg
    with llg.from_case(c, form="stencil") as h:
        h.mc_pt_setup(ladder, cone=0.6)
        E, M = h.mc_pt_sweeps(16, sampler, seed=11, sweep0=2, exchange_every=2, measure_every=4)
        st = h.mc_pt_state()
        ms = [h.mc_pt_get_state(q) for q in range(len(ladder))]
    assert list(st["image_of_rung"]) == pt["image"]
    assert np.array_equal(st["swaps_accepted"], pt["acc"][:5]) and np.array_equal(st["swaps_attempted"], pt["att"][:5])
    assert np.all(st["swaps_attempted"][::2] == 4) and np.all(st["swaps_attempted"][1::2] == 4)
    if setup == "TRIANGULAR_J3DKB":   # both outcomes occur on this ladder
        assert 0 < np.sum(st["swaps_accepted"]) < np.sum(st["swaps_attempted"])
    for q in range(len(ladder)):
        assert np.max(np.abs(ms[q] - pt["m"][pt["image"][q]])) <= 1e-12
    for r in range(len(ladder)):
        assert st["replicas"][r][1] == pt["st"][r][2] and abs(st["replicas"][r][0] - pt["st"][r][0]) <= 1e-12
    assert np.max(np.abs(E - E_ref)) <= 1e-12 * np.max(np.abs(E_ref))
    assert np.max(np.abs(M - M_ref)) <= 1e-9


def test_pt_reference_ptmc_run(oracle_lib):
    """The reference's tests/serial/PTMC is a parallel-tempering run: its 40
    temperatures as one ladder of replicas exchanging every 4 sweeps; energy per
    site by rung within the run's error bars."""
    from matjes_amd import llg
    ref = np.loadtxt(os.path.join(GOLDEN, "legacy_ptmc_ising20.txt"))
    c = _lattice_case(20, 20, (True, True, False), J=0.5e-3, K=5e-7, moment=3.0)
    n = c["lat"].ncell
    ladder = ref[:, 0] * oracle_lib.K_B
    with llg.from_case(c, form="stencil") as h:
        h.mc_pt_setup(ladder, cone=0.5)
        h.mc_pt_sweeps(500, ISING, seed=5, exchange_every=4)
        E, _ = h.mc_pt_sweeps(6000, ISING, seed=5, sweep0=500, exchange_every=4, measure_every=4)
        st = h.mc_pt_state()
    assert sorted(st["image_of_rung"]) == list(range(len(ladder)))
    assert np.all(st["swaps_accepted"] > 0)
    for q, (T, e_ref, e_err, _, _) in enumerate(ref):
        e, se = _mean_err(E[:, q] / n)
        assert abs(e - e_ref) <= 4.0 * np.hypot(e_err, se) + 1e-9, (T, e, se, e_ref, e_err)
