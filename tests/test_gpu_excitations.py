"""Time-dependent fields on the device (mj_set_bfield_maps / mj_steps_driven)
against the oracle's per-stage field and the static-field path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EULER, HEUN, SIB, DEPONDT = 1, 2, 5, 6


def _case(seed=3, size=(16, 16, 1)):
    from matjes_amd import workloads
    return workloads.build_case(dict(workloads.TRIANGULAR_J3DKB, size=list(size), seed=seed))


@pytest.mark.parametrize("integ,T,size", [(HEUN, 0.0, 16), (HEUN, 10.0, 16), (EULER, 0.0, 16), (SIB, 10.0, 16),
                                          (DEPONDT, 0.0, 16), (HEUN, 10.0, 64)])
def test_driven_maps_vs_oracle(oracle_lib, integ, T, size):
    """A plane-wave-like excitation (three maps, phasor weights) on top of the
    exchange/DMI/anisotropy terms: trajectory and energies against the oracle
    (64x64: the per-Hamiltonian compiled kernels)."""
    from matjes_amd import excitations as X, llg
    c = _case(size=(size, size, 1))
    n = c["lat"].ncell
    rng = np.random.default_rng(5)
    A, psi, off = rng.uniform(0.5, 1.5, n), rng.uniform(-3, 3, n), rng.normal(scale=0.3, size=(n, 3))
    nsteps, dt = 40, 1.0
    times = X.stage_times(nsteps, integ, dt)
    maps, w = X.map_weights(X.Phasor(3.0, 30.0, [0.5, -1.0, 2.0], 12.0, 0.25), A, psi, off, times)
    kT = T * oracle_lib.K_B
    with llg.from_case(c, form="stencil") as h:
        h.set_bfield_maps(maps)
        E = h.steps_driven(nsteps, integ, dt, 0.1, w, kT=kT, seed=4, energies=True)
        m = h.get_state()
    b_stage = X.field_of(maps, w)
    ref, E_ref = oracle_lib.Oracle(c["H"]).steps(c["m0"], nsteps, integ, dt, 0.1, kT=kT, seed=4, energies=True,
                                                 b_stage=b_stage)
    assert np.max(np.abs(m - ref)) <= 1e-12
    assert np.max(np.abs(E - E_ref)) <= 1e-12 * np.max(np.abs(E_ref))


@pytest.mark.parametrize("integ", [HEUN, SIB])
def test_driven_uniform_switch_on_equals_static_field(integ):
    """A uniform field switched on at step n_on (no maps): zero field before, then
    exactly the static-field trajectory (same Zeeman arithmetic), bit for bit."""
    from matjes_amd import excitations as X, llg, workloads
    c = workloads.build_case(dict(workloads.SQUARE_JDKB, size=[24, 20, 1], seed=6, B=[0.0, 0.0, 0.0]))
    Bvec = np.array([0.0, 0.5, 2.0])
    nsteps, n_on, dt = 30, 9, 1.0
    times = X.stage_times(nsteps, integ, dt)
    w = X.uniform_weights(X.Heaviside(times[n_on, 0] - 1e-9, 1e9, Bvec, Bvec), times)
    c["H"].add_zeeman(np.zeros(3))
    with llg.from_case(c, form="stencil") as h:
        h.set_bfield_maps(np.zeros((0, c["lat"].ncell, 3)))
        h.steps_driven(n_on, integ, dt, 0.1, w[:n_on])
        m_on = h.get_state()
        h.steps_driven(nsteps - n_on, integ, dt, 0.1, w[n_on:], step0=n_on)
        m = h.get_state()
    c["H"].add_zeeman(Bvec)
    with llg.from_case(dict(c, m0=m_on), form="stencil") as h:
        h.steps(nsteps - n_on, integ, dt, 0.1, step0=n_on)
        ref = h.get_state()
    assert np.array_equal(m, ref)


def test_driven_errors():
    from matjes_amd import _lib, llg, workloads
    c = workloads.build_case(dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=1, B=[0.0, 0.0, 0.0]))
    with llg.from_case(c, form="stencil") as h:
        with pytest.raises(_lib.MatjesError):      # no Zeeman term
            h.steps_driven(2, HEUN, 1.0, 0.1, np.zeros((2, 2, 1, 3)))
        with pytest.raises(_lib.MatjesError):      # at most 4 maps
            h.set_bfield_maps(np.zeros((5, c["lat"].ncell, 3)))


@pytest.mark.parametrize("integ", [EULER, HEUN, SIB, DEPONDT])
def test_per_step_temperature_vs_oracle(oracle_lib, integ):
    """mj_steps_temperature with a temperature pulse (TPulse.f90's
    exponential form, T0 = 0: the first steps carry no thermal term): the
    oracle gets the same Philox normals scaled per step as the reference forms
    the amplitude (sqrt(alpha kT_j hbar / dt_stage); SIB and Depondt 2 alpha
    kT_j hbar / dt); states and energies to 1e-11.  A constant schedule gives
    mj_steps' bits."""
    from matjes_amd import excitations, llg
    c = _case(seed=4)
    n, ts, alpha, seed, step0 = 60, 0.5, 0.05, 21, 0
    pulse = excitations.TPulse(T0=0.0, I_exp=40.0)
    kts = np.array([pulse.kT((step0 + j) * ts) for j in range(n)])
    assert kts[0] == 0.0 and kts.max() > 30.0 * oracle_lib.K_B
    dt = ts / 2.0 if integ == HEUN else ts
    f = 1.0 if integ in (EULER, HEUN) else 2.0
    noise = np.stack([oracle_lib.thermal_field(seed, step0 + j, c["lat"].ncell,
                                               np.sqrt(f * alpha * kts[j] * oracle_lib.HBAR / dt))
                      if kts[j] * alpha > 1e-10 else np.zeros((c["lat"].ncell, 3)) for j in range(n)])
    o = oracle_lib.Oracle(c["H"])
    mo, Eo = o.steps(c["m0"], n, integ, ts, alpha, kT=kts.max(), noise=noise, energies=True)
    with llg.from_case(c, form="stencil") as h:
        Eg = h.steps_temperature(kts, integ, ts, alpha, seed=seed, step0=step0, energies=True)
        mg = h.get_state()
        assert np.max(np.abs(mg - mo)) <= 1e-11
        assert np.max(np.abs(Eg - Eo)) <= 1e-11 * np.max(np.abs(Eo))
        kT = 12.0 * oracle_lib.K_B
        h.set_state(c["m0"])
        E1 = h.steps_temperature(np.full(70, kT), integ, ts, alpha, seed=seed, step0=step0, energies=True)
        m1 = h.get_state()
        h.set_state(c["m0"])
        E2 = h.steps(70, integ, ts, alpha, kT=kT, seed=seed, step0=step0, energies=True)
        assert np.array_equal(m1, h.get_state()) and np.array_equal(E1, E2)


def test_spindynamics_with_a_temperature_pulse(tmp_path):
    """llg.spindynamics(temperature=pulse): kT of step j from the pulse at the
    step's start time; the same state as one mj_steps_temperature call over
    the schedule, and EM.dat's temperature column the pulse's."""
    from matjes_amd import emio, excitations, llg
    c = _case(seed=5)
    pulse = excitations.TPulse(T0=5.0, I_exp=40.0)
    ts, n = 0.5, 41
    with llg.from_case(c, form="stencil") as h:
        llg.spindynamics(h, n, HEUN, ts, 0.05, efreq=10, seed=8, outdir=str(tmp_path), temperature=pulse)
        m_loop = h.get_state()
        h.set_state(c["m0"])
        h.steps_temperature([pulse.kT(j * ts) for j in range(n)], HEUN, ts, 0.05, seed=8)
        assert np.array_equal(m_loop, h.get_state())
    em = np.loadtxt(tmp_path / "EM.dat")                  # column 15: T (K)
    steps = em[:, 0].astype(int)
    assert list(steps) == [1, 11, 21, 31, 41] and emio.EM_HEADER[14].startswith("15:T")
    np.testing.assert_allclose(em[:, 14], [pulse.temperature((j - 1) * ts) for j in steps], rtol=1e-11)


@pytest.mark.parametrize("shape_r", ["spot gaussian 6.0 5.0 0.0 2.5", "beam laguerre-gauss-By 8 8 0 1 1 6.0 1.5"])
def test_spatial_shape_excitation_vs_oracle(oracle_lib, shape_r):
    """A time shape read from the reference's input form (EMwave) on a
    real-space shape (a Gaussian spot, the B_y component of a Laguerre-Gauss
    beam) at the lattice's site positions: the device trajectory against the
    oracle's with the same per-stage field, 1e-12."""
    from matjes_amd import excitations as X, llg
    c = _case()
    R = c["lat"].cell_positions()
    shape_t = X.read_shape("Bfield EMwave 0 200 3  2.0 -1.0 0.5  9.0 9.0 9.0  0.5 0.0 0.25  20.0 8.0")
    nsteps, dt = 40, 1.0
    times = X.stage_times(nsteps, HEUN, dt)
    maps, w = X.shape_maps(shape_t, X.ShapeR(shape_r), R, times)
    assert np.max(np.abs(X.field_of(maps, w))) > 0.1
    with llg.from_case(c, form="stencil") as h:
        h.set_bfield_maps(maps)
        E = h.steps_driven(nsteps, HEUN, dt, 0.1, w, energies=True)
        m = h.get_state()
    ref, E_ref = oracle_lib.Oracle(c["H"]).steps(c["m0"], nsteps, HEUN, dt, 0.1, energies=True,
                                                 b_stage=X.field_of(maps, w))
    assert np.max(np.abs(m - ref)) <= 1e-12
    assert np.max(np.abs(E - E_ref)) <= 1e-12 * np.max(np.abs(E_ref))
