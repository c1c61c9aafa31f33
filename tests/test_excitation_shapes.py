"""Time shapes of the excitations (src/excitations/exc_t_*.f90), the
temperature pulse (src/excitations/TPulse.f90) and the real-space shapes
(src/excitations/exc_r.f90), pinned by values the formulas fix in closed
form: the Gaussian's 1/2 in its exponent against EMwave's none, the period
and pi-phase conversions of EMwave's input, the half-open windows and their
clipping to t0 -+ 10 tau, the ramp's midpoint, the algebraic power, the
pulse's peak, plateau and cooling law; the Laguerre polynomials at 0 and
low order, the beam's waist, azimuthal phase, Gouy phase and wavefront
curvature at the Rayleigh length, its azimuthal B in the waist plane, and
the strict boundaries of the square and ball shapes.  The reference's own
shapes need GSL (absent), so no reference output pins them."""
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


# ---- real-space shapes (exc_r.f90)

def test_laguerre_closed_forms():
    from math import comb
    x = np.array([0.0, 0.3, 1.7, 4.0])
    for n in range(5):
        for a in (0, 1, 2, 3):
            assert ex.laguerre(n, a, 0.0) == pytest.approx(comb(n + a, n), rel=1e-14)   # L_n^a(0)
    for a in (0, 1, 2):
        np.testing.assert_allclose(ex.laguerre(2, a, x), ((a + 1) * (a + 2) - 2 * (a + 2) * x + x * x) / 2.0,
                                   rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(ex.laguerre(3, 0, x), (-x ** 3 + 9 * x * x - 18 * x + 6) / 6.0, rtol=1e-13, atol=1e-13)


def test_square_cylinder_gaussian_shapes():
    R = np.array([[1.0, 2.0, 0.0], [3.9, 2.0, 0.0], [4.0, 2.0, 0.0], [3.0, 4.0, 0.0], [1.0, 2.0, 3.5]])
    A, psi = ex.ShapeR("s square 1 2 0 3").amplitude_phase(R)
    assert list(A) == [1, 1, 0, 1, 0] and not psi.any()                 # |R - c|_inf < 3, strict
    A, _ = ex.ShapeR("s cylinder 1 2 0 3").amplitude_phase(R)
    assert list(A) == [1, 1, 0, 1, 0]                                   # a ball: |(2, 2, 0)| = 2.83 < 3, z counts
    A, _ = ex.ShapeR("s gaussian 1 2 0 2").amplitude_phase(R)
    np.testing.assert_allclose(A[[0, 2]], [1.0, np.exp(-9.0 / 4.0)], rtol=1e-15)
    A, _ = ex.ShapeR("s gaussian 0 0 0 0.01").amplitude_phase(np.array([[1.0, 0, 0]]))
    assert A[0] == np.exp(-200.0)                                       # exponent floor
    A, psi = ex.ShapeR("s plane").amplitude_phase(R)
    assert np.array_equal(A, np.ones(5)) and not psi.any()


def test_laguerre_gauss_beam_values():
    w0, period = 50.0, 2.0
    lam = ex.C_LIGHT * period
    b = ex.ShapeR("b laguerre-gauss 0 0 0 0 0 %r %r" % (w0, period))
    A, psi = b.amplitude_phase(np.array([[w0, 0.0, 0.0]]))              # l = p = 0, waist plane, r = w0
    assert A[0] == pytest.approx(np.sqrt(2.0 / np.pi) / w0 * np.exp(-1.0), rel=1e-14) and psi[0] == 0.0
    b1 = ex.ShapeR("b laguerre-gauss 0 0 0 1 0 %r %r" % (w0, period))   # l = 1: phase = azimuth at z = 0
    pts = np.array([[10.0, 0, 0], [0, 10.0, 0], [-10.0, 0, 0], [7.0, -7.0, 0]])
    A, psi = b1.amplitude_phase(pts)
    np.testing.assert_allclose(psi, np.arctan2(pts[:, 1], pts[:, 0]), rtol=0, atol=1e-15)
    r = np.hypot(pts[:, 0], pts[:, 1])
    np.testing.assert_allclose(A, np.sqrt(4.0 / np.pi) / w0 * (r * np.sqrt(2.0) / w0) * np.exp(-r ** 2 / w0 ** 2),
                               rtol=1e-14)
    # at the Rayleigh length: width w0 sqrt 2, Gouy phase pi/4, wavefront radius 2 z_R
    z_R = np.pi * w0 ** 2 / lam
    rr = 20.0
    A, psi = b.amplitude_phase(np.array([[rr, 0.0, z_R]]))
    w = w0 * np.sqrt(2.0)
    assert A[0] == pytest.approx(np.sqrt(2.0 / np.pi) / w * np.exp(-rr ** 2 / w ** 2), rel=1e-13)
    assert psi[0] == pytest.approx(2 * np.pi * z_R / lam - np.pi / 4 + np.pi * rr ** 2 / (lam * 2 * z_R), rel=1e-13)


def test_laguerre_gauss_field_components():
    """l = p = 0 in the waist plane: B is azimuthal, 4 N r / (omega w0^3)
    exp(-r^2/w0^2) (times 1e3: V fs / nm^2 to T), a quarter period behind."""
    w0, period = 50.0, 2.0
    omega = 2 * np.pi / period
    bx = ex.ShapeR("b laguerre-gauss-Bx 0 0 0 0 0 %r %r" % (w0, period))
    by = ex.ShapeR("b laguerre-gauss-By 0 0 0 0 0 %r %r" % (w0, period))
    r = 10.0
    mag = 1e3 * 4.0 / omega * np.sqrt(2.0 / np.pi) * r / w0 ** 3 * np.exp(-r ** 2 / w0 ** 2)
    Ay, py = by.amplitude_phase(np.array([[r, 0.0, 0.0]]))             # on +x: along +y
    assert Ay[0] == pytest.approx(mag, rel=1e-13) and py[0] == pytest.approx(-np.pi / 2, rel=1e-15)
    Ax, _ = bx.amplitude_phase(np.array([[r, 0.0, 0.0]]))
    assert Ax[0] <= 1e-12 * mag
    Ax, px = bx.amplitude_phase(np.array([[0.0, r, 0.0]]))             # on +y: along -x
    assert Ax[0] == pytest.approx(mag, rel=1e-13) and px[0] == pytest.approx(np.pi / 2, rel=1e-15)


def test_file_shape_and_maps(tmp_path):
    """The file shape contributes its offsets only; the maps and weights of
    shape_maps give update_exc's b_i(t) = T(t) (A_i cos(phi(t) + psi_i) + o_i)."""
    lat = __import__("matjes_amd.lattice", fromlist=["x"]).square_film(6, 5, 1)
    R = lat.cell_positions()
    off = np.random.default_rng(3).normal(size=(len(R), 3))
    np.savetxt(tmp_path / "Bexc_file", off)
    f = ex.ShapeR("Bexc file", directory=str(tmp_path))
    A, psi = f.amplitude_phase(R)
    assert not A.any() and np.array_equal(f.offset, off)
    shape_t = ex.read_shape("Bfield EMwave 0 100 3  1.0 0.5 -2.0  8.0 8.0 8.0  0.25 0.0 1.0  30.0 12.0")
    spot = ex.ShapeR("spot gaussian 2 2 0 1.5")
    times = np.array([[20.0, 20.5], [31.0, 31.5]])
    for shape_r in (spot, f):
        maps, w = ex.shape_maps(shape_t, shape_r, R, times)
        b = ex.field_of(maps, w)
        A, psi = shape_r.amplitude_phase(R)
        o = shape_r.offset if shape_r.offset is not None else np.zeros((len(R), 3))
        for idx in np.ndindex(times.shape):
            T, phi = shape_t(times[idx])
            direct = T[None, :] * (A[:, None] * np.cos(phi + psi)[:, None] + o)
            np.testing.assert_allclose(b[idx], direct, rtol=1e-13, atol=1e-14)
