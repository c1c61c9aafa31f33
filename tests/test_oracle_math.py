"""Pins of the oracle that do not come from reference fixtures (CPU):
published known-answer vectors, closed forms, the reference's legacy golden
file, invariants and convergence orders of the integrators."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from matjes_amd import lattice as L
from matjes_amd import workloads


# Random123 known-answer vectors for philox4x32-10 (Salmon et al., SC'11)
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_kat(oracle_lib, ctr, key, out):
    assert oracle_lib.philox(ctr, key) == out


def test_normals_mapping(oracle_lib):
    """Box-Muller mapping of the Philox words, recomputed independently."""
    seed, step, site = 0x1234567890ABCDEF, 77, 12345
    g = oracle_lib.thermal_field(seed, step, site + 1, 1.0)[site]
    key = [seed & 0xFFFFFFFF, seed >> 32]
    x = oracle_lib.philox([site, step, 0, 0], key)
    r1 = np.sqrt(-2.0 * np.log((x[0] + 1) * 2.0 ** -32))
    r2 = np.sqrt(-2.0 * np.log((x[2] + 1) * 2.0 ** -32))
    a1, a2 = 2 * np.pi * x[1] * 2.0 ** -32, 2 * np.pi * x[3] * 2.0 ** -32
    want = [r1 * np.cos(a1), r1 * np.sin(a1), r2 * np.cos(a2)]
    np.testing.assert_allclose(g, want, rtol=1e-14, atol=1e-15)


def test_normals_statistics(oracle_lib):
    n = 200000
    g = oracle_lib.thermal_field(5, 3, n, 1.0)
    assert abs(g.mean()) < 0.01
    assert abs(g.var() - 1.0) < 0.01
    assert abs(np.mean(g[:, 0] * g[:, 1])) < 0.01


def _single_spin(B, moment=1.0):
    lat = L.Lattice(np.eye(3), (1.0, 1.0, 1.0), (1, 1, 1), (False, False, False), [("Fe", (0, 0, 0), moment)])
    H = L.Hamiltonian(lat)
    H.add_zeeman(B)
    return H


def test_euler_precession_closed_form(oracle_lib):
    """Euler at alpha=0 rotates by atan(mu_B mu_s B dt / hbar) per step."""
    H = _single_spin([0.0, 0.0, 10.0])
    o = oracle_lib.Oracle(H)
    m = np.array([[1.0, 0.0, 0.0]])
    th = np.arctan(L.MU_B * 10.0 * 1.0 / L.HBAR)
    for n in range(1, 51):
        m = o.steps(m, 1, 1, 1.0, 0.0)
        assert abs(m[0, 1] - np.sin(n * th)) < 2e-15
        assert abs(m[0, 0] - np.cos(n * th)) < 2e-15


def test_euler_matches_reference_legacy_file(oracle_lib):
    """tests/serial/isolated_spin_precession/Euler/REF (50 steps, 12 digits): the
    reference's own data is a rotation by a fixed angle per step; the oracle's
    Euler reproduces every row once the field matches that angle."""
    d = np.loadtxt(os.path.join(GOLDEN, "legacy_precession_euler.txt"))
    th = np.arcsin(d[0, 2])
    n = np.arange(1, 51)
    assert np.max(np.abs(d[:, 2] - np.sin(n * th))) < 1e-11
    assert np.max(np.abs(d[:, 1] - np.cos(n * th))) < 1e-11
    B = np.tan(th) * L.HBAR / (L.MU_B * 1.0)
    o = oracle_lib.Oracle(_single_spin([0.0, 0.0, B]))
    m = np.array([[1.0, 0.0, 0.0]])
    for k in range(50):
        m = o.steps(m, 1, 1, 1.0, 0.0)
        assert abs(m[0, 1] - d[k, 2]) < 1e-11 and abs(m[0, 0] - d[k, 1]) < 1e-11


def test_two_spin_dipolar_closed_form(oracle_lib):
    """B_1 = dip_pref/(4 pi) mu^2 (3 r(r.m2) - m2)/|r|^3 for two moments on an open line."""
    a = 0.4
    lat = L.Lattice(np.eye(3), (a, a, a), (2, 1, 1), (False, False, False), [("Fe", (0, 0, 0), 2.0)])
    H = L.Hamiltonian(lat)
    H.add_dipolar()
    o = oracle_lib.Oracle(H)
    m = workloads.random_sphere(2, 3)
    B = o.field(m)
    r = np.array([1.0, 0.0, 0.0])
    for i, j in ((0, 1), (1, 0)):
        want = L.DIP_PREF / (4 * L.PI) * 4.0 * (3 * r * np.dot(r, m[j]) - m[j]) / a ** 3
        np.testing.assert_allclose(B[i], want, rtol=1e-13, atol=1e-20)


@pytest.mark.parametrize("periodic", [(False, False, False), (True, True, False)])
def test_fft_convolution_is_circular_sum(oracle_lib, periodic):
    s = dict(workloads.FILM_DIP, size=[4, 3, 2], periodic=list(periodic), seed=5)
    c = workloads.build_case(s)
    dc = oracle_lib.DipolarConv(c["H"].dipolar, c["lat"].dim_lat)
    h1 = dc.h(c["m0"])
    h2 = dc.h_direct(c["m0"])
    assert np.max(np.abs(h1 - h2)) <= 1e-13 * np.max(np.abs(h2))


def test_builder_shells_and_dmi():
    sq = workloads.make_lattice(dict(workloads.SQUARE_JDKB, size=[8, 8, 1]))
    tri = workloads.make_lattice(dict(workloads.TRIANGULAR_J3DKB, size=[8, 8, 1]))
    for lat, counts in ((sq, (4, 4, 4)), (tri, (6, 6, 6))):
        for sh, cnt in zip((1, 2, 3), counts):
            offs, _ = L.shell_offsets(lat, sh)
            assert len(offs) == cnt
        for off in L.shell_offsets(lat, 1)[0]:
            r = off[0] * lat.areal[0] + off[1] * lat.areal[1] + off[2] * lat.areal[2]
            want = np.cross([0.0, 0.0, 1.0], r / np.linalg.norm(r))   # interfacial: z x r for a mediator below
            np.testing.assert_allclose(L.dmi_vector(lat, off, "Ir"), want, atol=1e-15)


def _conv_error(orc, integ, dts, T_total=100.0):
    s = dict(workloads.SQUARE_JDKB, size=[6, 6, 1], seed=9)
    c = workloads.build_case(s)
    o = orc.Oracle(c["H"])
    ref = o.steps(c["m0"], int(round(T_total / 0.0125)), integ, 0.0125, 0.1)
    return [np.max(np.abs(o.steps(c["m0"], int(round(T_total / dt)), integ, dt, 0.1) - ref)) for dt in dts]


@pytest.mark.parametrize("integ", [5, 6])
def test_sib_depondt_second_order(oracle_lib, integ):
    e1, e2 = _conv_error(oracle_lib, integ, [1.0, 0.5])
    assert 3.0 < e1 / e2 < 5.0


def test_reference_heun_is_first_order(oracle_lib):
    """The reference's integration 2 (second stage from the predictor) is first order."""
    e1, e2 = _conv_error(oracle_lib, 2, [1.0, 0.5])
    assert 1.6 < e1 / e2 < 2.4


@pytest.mark.parametrize("integ", [5, 6])
def test_norm_conservation_and_energy_descent(oracle_lib, integ):
    s = dict(workloads.TRIANGULAR_J3DKB, size=[8, 8, 1], seed=4)
    c = workloads.build_case(s)
    o = oracle_lib.Oracle(c["H"])
    m, E = o.steps(c["m0"], 400, integ, 1.0, 0.1, energies=True)
    assert np.max(np.abs(np.linalg.norm(m, axis=1) - 1.0)) < 1e-13
    assert np.all(np.diff(E) <= 1e-12 * np.abs(E[:-1]))


def test_reciprocal_division_is_correctly_rounded(oracle_lib):
    """The device divides as q0 = a*RN(1/b) plus one fma correction (Markstein);
    it must equal IEEE a/b, which the reference computes."""
    assert oracle_lib.lib().or_div_rn_mismatches(20000000, 12345) == 0


@pytest.mark.parametrize("integ", [1, 2])
def test_update_time_single_spin_closed_form(oracle_lib, integ):
    """update_time (update_time.f90:41-75): one spin in B along z has |B_eff| = mu_s mu_B B,
    so every step after the first runs with 2 pi hbar / (mu_s mu_B B) / discretization,
    and Euler/Heun rotate by the closed-form angle for that step."""
    B, disc = 7.0, 15.0
    o = oracle_lib.Oracle(_single_spin([0.0, 0.0, B]))
    ts = np.zeros(12)
    m = o.steps(np.array([[1.0, 0.0, 0.0]]), 12, integ, 0.3, 0.0, discretization=disc, timesteps=ts)
    want = 2.0 * L.PI * L.HBAR / (L.MU_B * B) / disc
    assert ts[0] == 0.3
    np.testing.assert_allclose(ts[1:], want, rtol=1e-15)
    # plain steps with the same timesteps give the same state
    m2 = np.array([[1.0, 0.0, 0.0]])
    for t in ts:
        m2 = o.steps(m2, 1, integ, t, 0.0)
    assert np.max(np.abs(m - m2)) == 0.0
    if integ == 1:
        ang = np.arctan(L.MU_B * B * 0.3 / L.HBAR) + 11 * np.arctan(L.MU_B * B * want / L.HBAR)
        np.testing.assert_allclose(m[0], [np.cos(ang), np.sin(ang), 0.0], atol=1e-14)


def test_update_time_zero_field_stops(oracle_lib):
    o = oracle_lib.Oracle(_single_spin([0.0, 0.0, 0.0]))
    with pytest.raises(RuntimeError):
        o.steps(np.array([[1.0, 0.0, 0.0]]), 2, 1, 0.3, 0.0, discretization=15.0, timesteps=np.zeros(2))
