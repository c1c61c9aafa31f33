"""Time shapes of the excitations (src/excitations/exc_t_*.f90) and the
temperature pulse (src/excitations/TPulse.f90), pinned by values the
formulas fix in closed form: the Gaussian's 1/2 in its exponent against
EMwave's none, the period and pi-phase conversions of EMwave's input, the
half-open windows and their clipping to t0 -+ 10 tau, the ramp's midpoint,
the algebraic power, and the pulse's peak, plateau and cooling law."""
import numpy as np
import pytest

from matjes_amd import excitations as ex
from matjes_amd.lattice import K_B


def test_gauss_values_and_window():
    g = ex.read_shape("Bfield gauss 0.0 1000.0 3  2.0 -1.0 0.5  40.0 4.0")
    b, _ = g(40.0)
    assert np.array_equal(b, [2.0, -1.0, 0.5])
    b, _ = g(44.0)                                          # one tau: e^(-1/2)
    np.testing.assert_allclose(b, np.array([2.0, -1.0, 0.5]) * np.exp(-0.5), rtol=1e-15)
    assert g.t_start == 0.0 and g.t_end == 80.0            # clipped to t0 + 10 tau
    assert np.array_equal(g(80.0)[0], np.zeros(3)) and np.any(g(79.99)[0] != 0)
    assert np.array_equal(ex.read_shape("B gauss 30 1000 3 1 1 1 40 4")(29.9)[0], np.zeros(3))


def test_rampe_values():
    r = ex.read_shape("Bfield rampe 10.0 30.0 3  0.0 1.0 -2.0  4.0 3.0 2.0")
    assert np.array_equal(r(10.0)[0], [0.0, 1.0, -2.0])     # t_start included
    np.testing.assert_allclose(r(20.0)[0], [2.0, 2.0, 0.0], rtol=0, atol=1e-15)
    assert np.array_equal(r(30.0)[0], np.zeros(3))          # t_end excluded
    assert np.array_equal(r(9.99)[0], np.zeros(3))


def test_emwave_input_conversions():
    # period 10 fs, phase 0.5 pi, envelope exp(-((t - 50)/5)^2) (no 1/2)
    w = ex.read_shape("Bfield EMwave 0.0 1000.0 3  1.0 2.0 0.0  10.0 10.0 10.0  0.5 0.5 0.5  50.0 5.0")
    assert np.allclose(w.omega, 2.0 * np.pi / 10.0) and np.allclose(w.phi, 0.5 * np.pi)
    b, _ = w(50.0)                                          # cos(10 pi + pi/2) = 0
    assert np.max(np.abs(b)) <= 1e-14
    b, _ = w(52.5)                                          # cos(11 pi) = -1, envelope e^(-1/4)
    np.testing.assert_allclose(b, [-np.exp(-0.25), -2.0 * np.exp(-0.25), 0.0], rtol=1e-13, atol=1e-15)
    b, _ = w(55.0)                                          # cos(11.5 pi) = 0
    assert np.max(np.abs(b)) <= 1e-14
    assert w.t_start == 0.0 and w.t_end == 100.0
    # the window ends at t0 + 10 tau (so the exponent stays above -100 and the
    # reference's guard at -200 never acts)
    narrow = ex.EMwave(0.0, 1000.0, [1.0, 1.0, 1.0], [0.1] * 3, [0.0] * 3, 50.0, 0.1)
    assert np.array_equal(narrow(51.0)[0], np.zeros(3)) and np.all(narrow(50.99)[0] != 0)


def test_algebraic_power():
    a = ex.read_shape("Bfield algebraic 0.0 100.0 3  8.0 -4.0 1.0  -1.0 2.0")
    np.testing.assert_allclose(a(1.0)[0], [2.0, -1.0, 0.25], rtol=1e-15)   # I0 / (1 - (-1))^2
    np.testing.assert_allclose(a(3.0)[0], [0.5, -0.25, 0.0625], rtol=1e-15)
    assert np.array_equal(a(100.0)[0], np.zeros(3))


def test_read_shape_errors():
    with pytest.raises(ValueError):
        ex.read_shape("Bfield gauss 0 10 3 1 1 1 5")          # one real short
    with pytest.raises(ValueError):
        ex.read_shape("Bfield wobble 0 10 3 1 1 1 5 1")


def test_tpulse_exponential_form():
    p = ex.TPulse()                                        # get_parameter_TPulse's defaults
    assert p.temperature(22.57) == pytest.approx(150.0 + 137.448, rel=1e-15)
    sig = np.sqrt(52.29)
    assert p.temperature(22.57 + sig) == pytest.approx(150.0 + 137.448 * np.exp(-0.5), rel=1e-14)
    assert p.temperature(22.57 + 3.0 * sig) == 150.0 and p.temperature(0.0) == 150.0
    assert p.kT(22.57) == pytest.approx((150.0 + 137.448) * K_B, rel=1e-15)


def test_tpulse_cooling_form_and_scaling():
    p = ex.TPulse(forme="cool", alpha=2.0)
    # t_cut from the unscaled amplitudes, then I_exp and I_over_x doubled
    assert p.t_cut == pytest.approx((22.57 - 19.62) / 286.415 / 0.47 + 22.57, rel=1e-15)
    assert p.I_exp == 2.0 * 137.448 and p.I_over_x == 2.0 * 286.415
    t_peak = 25 + 22.57 - 1e-9                             # just before t_start + t_exp: the Gaussian rise
    assert p.temperature(t_peak) == pytest.approx(150.0 + 2.0 * 137.448, rel=1e-12)
    t = 25 + 22.57 + 0.01                                  # cooling: T0 + I_over_x / (t - t_start - t_frac)^expo
    assert p.temperature(t) == pytest.approx(150.0 + 2.0 * 286.415 / (t - 25 - 19.62) ** 0.47, rel=1e-14)
    assert p.temperature(25 + p.t_cut + 1.0) == 150.0
    assert p.temperature(25 + 22.57 - 3.0 * np.sqrt(52.29)) == 150.0
