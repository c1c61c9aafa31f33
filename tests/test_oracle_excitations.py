"""Time-dependent fields (the reference's excitations) on the CPU: the map and
weight decomposition against the reference's formula, and the oracle's
per-stage field against closed forms."""
import numpy as np
import pytest

EULER, HEUN, SIB, DEPONDT = 1, 2, 5, 6


def _paramagnet(nx=4, ny=4, moment=2.0):
    from matjes_amd import lattice as L
    lat = L.square_film(nx, ny, 1, alat=1.0, moment=moment, mediator_dz=None, periodic=(True, True, False))
    H = L.Hamiltonian(lat)
    H.add_zeeman(np.zeros(3))
    return lat, H


def test_decomposition_matches_update_exc():
    """b_i(t) = T(t) (x) (A_i cos(phi(t) + psi_i) + o_i) (excitations.f90:204-210) from
    three maps and their weights."""
    from matjes_amd import excitations as X
    rng = np.random.default_rng(0)
    n = 50
    A, psi, off = rng.uniform(0.5, 2.0, n), rng.uniform(-3, 3, n), rng.normal(size=(n, 3))
    shape = X.Phasor(0.0, 100.0, [0.3, -1.0, 2.0], 25.0, 0.4)
    times = X.stage_times(12, HEUN, 1.0)
    maps, w = X.map_weights(shape, A, psi, off, times)
    b = X.field_of(maps, w)
    for idx in np.ndindex(times.shape):
        T, phi = shape(times[idx])
        ref = T[None, :] * (A[:, None] * np.cos(phi + psi)[:, None] + off)
        assert np.max(np.abs(b[idx] - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_shapes_and_stage_times():
    from matjes_amd import excitations as X
    h = X.Heaviside(5.0, 10.0, [1, 0, 0], [0, 0, 2])
    assert np.array_equal(h(4.999)[0], [0, 0, 0]) and np.array_equal(h(5.0)[0], [1, 0, 0])
    assert np.array_equal(h(10.0)[0], [0, 0, 2])
    p = X.Phasor(0.0, 10.0, [1, 1, 1], 4.0, 0.5)
    assert abs(p(1.0)[1] - (-2 * np.pi / 4.0 + 0.5 * np.pi)) < 1e-15 and np.array_equal(p(11.0)[0], [0, 0, 0])
    assert np.allclose(X.stage_times(3, EULER, 2.0)[:, 0], [2.0, 4.0, 6.0])
    assert np.allclose(X.stage_times(3, HEUN, 2.0), [[1, 1], [3, 3], [5, 5]])


@pytest.mark.parametrize("integ", [EULER, HEUN])
def test_switch_on_is_the_static_field_delayed(oracle_lib, integ):
    """A field switched on at step n_on: before it the spins do not move, after it
    the trajectory is the static-field one shifted by n_on steps, bit for bit."""
    from matjes_amd import excitations as X
    lat, H = _paramagnet()
    rng = np.random.default_rng(1)
    m0 = rng.normal(size=(lat.ncell, 3))
    m0 /= np.linalg.norm(m0, axis=1)[:, None]
    Bvec = np.array([0.0, 1.0, 3.0])
    nsteps, n_on, dt = 30, 7, 1.0
    times = X.stage_times(nsteps, integ, dt)
    t_on = times[n_on, 0] - 1e-9
    w = X.uniform_weights(X.Heaviside(t_on, 1e9, Bvec, Bvec), times)
    b_stage = X.field_of(np.ones((1, lat.ncell, 3)), w)
    o = oracle_lib.Oracle(H)
    m = o.steps(m0, nsteps, integ, dt, 0.05, b_stage=b_stage)
    H.add_zeeman(Bvec)
    o2 = oracle_lib.Oracle(H)
    ref = o2.steps(m0, nsteps - n_on, integ, dt, 0.05)
    assert np.array_equal(m, ref)
    m_before = o.steps(m0, n_on, integ, dt, 0.05, b_stage=b_stage[:n_on])
    assert np.array_equal(m_before, np.where(np.abs(m0) < 1e-50, 0.0, m0))


def test_constant_maps_are_the_static_map(oracle_lib):
    from matjes_amd import excitations as X, workloads
    c = workloads.build_case(dict(workloads.SQUARE_JDKB, size=[8, 8, 1], bmap_sigma=0.5, seed=2))
    o = oracle_lib.Oracle(c["H"])
    ref, E_ref = o.steps(c["m0"], 20, HEUN, 1.0, 0.1, energies=True)
    b = c["H"].b_field_sites()
    w = np.ones((20, 2, 1, 3))
    m, E = o.steps(c["m0"], 20, HEUN, 1.0, 0.1, energies=True, b_stage=X.field_of(b[None], w))
    assert np.array_equal(m, ref) and np.array_equal(E, E_ref)
