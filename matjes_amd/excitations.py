"""Time-dependent B order parameter (the reference's excitations) for
llg.Hamiltonian.steps_driven.

The reference sets the excited order parameter before every stage of a step
(update_exc(real_time + dt, ...), src/dynamics/spindynamics.f90:285-292) to

    b_i(t) = sum_e T_e(t) (x) ( A_e,i cos(phi_e(t) + psi_e,i) + o_e,i )

(src/excitations/excitations.f90:164-215): a time shape T_e, phi_e
(exc_t_*.f90) times a real-space shape A_e,i, psi_e,i, o_e,i (exc_r.f90).
Expanding the cosine makes b a sum of fixed site maps with time-dependent
weights,

    A cos psi (x) T cos phi  -  A sin psi (x) T sin phi  +  o (x) T,

which is what the device evaluates per stage (at most 4 maps).  Stage times
follow the reference: dt_stage = Butcher(1, l+1) * timestep / N_loop, so both
stages of the two-half-step Heun and the single Euler stage are at
real_time + timestep / N_loop; SIB and Depondt (no reference) use the
midpoint real_time + timestep / 2 for both stages.  The energy after a step
uses the field of its last stage (mag_lattice was updated last at that time).
"""
import math
import os

import numpy as np

PI = 3.14159265358979323846264338327950288


def stage_times(nsteps, integrator, timestep, t0=0.0):
    """(nsteps, stages) times at which the field of each stage is evaluated."""
    j = np.arange(nsteps, dtype=np.float64)
    if integrator == 1:           # Euler: N_loop = 1
        return (t0 + j * timestep + timestep)[:, None]
    if integrator == 2:           # two half steps of timestep / 2 each
        t = t0 + j * timestep + timestep / 2.0
        return np.stack([t, t], axis=1)
    t = t0 + j * timestep + timestep / 2.0   # SIB / Depondt: midpoint
    return np.stack([t, t], axis=1)


class Heaviside:
    """exc_t_heaviside.f90:53-62: val_start in [t_start, t_end), val_end after."""

    def __init__(self, t_start, t_end, val_start, val_end):
        self.t_start, self.t_end = float(t_start), float(t_end)
        self.v0 = np.asarray(val_start, dtype=np.float64)
        self.v1 = np.asarray(val_end, dtype=np.float64)

    def __call__(self, t):
        val = np.zeros(3)
        if self.t_start <= t < self.t_end:
            val = self.v0.copy()
        if t >= self.t_end:
            val = self.v1.copy()
        return val, 0.0


class Phasor:
    """exc_t_phasor.f90:47-62: I0 with phase -2 pi t / period + phi pi inside
    [t_start, t_end], zero outside."""

    def __init__(self, t_start, t_end, I0, period, phi):
        self.t_start, self.t_end = float(t_start), float(t_end)
        self.I0 = np.asarray(I0, dtype=np.float64)
        self.period, self.phi = float(period), float(phi)

    def __call__(self, t):
        if t < self.t_start or t > self.t_end:
            return np.zeros(3), 0.0
        return self.I0.copy(), -2.0 * PI / self.period * t + self.phi * PI


class Constant:
    """exc_t_const.f90: val inside [t_start, t_end), zero outside."""

    def __init__(self, t_start, t_end, val):
        self.t_start, self.t_end = float(t_start), float(t_end)
        self.v = np.asarray(val, dtype=np.float64)

    def __call__(self, t):
        if self.t_start <= t < self.t_end:
            return self.v.copy(), 0.0
        return np.zeros(3), 0.0


class Gauss:
    """exc_t_gauss.f90: I0 exp(-((t - t0)/tau)^2 / 2) inside [t_start, t_end),
    the interval first clipped to t0 -+ 10 tau (gauss_read_string)."""

    def __init__(self, t_start, t_end, I0, t0, tau):
        self.t0, self.tau = float(t0), float(tau)
        self.t_start = max(float(t_start), self.t0 - 10.0 * self.tau)
        self.t_end = min(float(t_end), self.t0 + 10.0 * self.tau)
        self.I0 = np.asarray(I0, dtype=np.float64)

    def __call__(self, t):
        if not (self.t_start <= t < self.t_end):
            return np.zeros(3), 0.0
        x = (t - self.t0) / self.tau
        return self.I0 * np.exp(-0.5 * x * x), 0.0


class Rampe:
    """exc_t_rampe.f90: linear from val_start at t_start towards val_end at
    t_end, inside [t_start, t_end), zero outside."""

    def __init__(self, t_start, t_end, val_start, val_end):
        self.t_start, self.t_end = float(t_start), float(t_end)
        self.v0 = np.asarray(val_start, dtype=np.float64)
        self.v1 = np.asarray(val_end, dtype=np.float64)

    def __call__(self, t):
        if not (self.t_start <= t < self.t_end):
            return np.zeros(3), 0.0
        return self.v0 + (self.v1 - self.v0) * (t - self.t_start) / (self.t_end - self.t_start), 0.0


class EMwave:
    """exc_t_EMwave.f90: I0 cos(omega t + phi) exp(-((t - t0)/tau)^2) inside
    [t_start, t_end) (clipped to t0 -+ 10 tau), zero where the exponent is
    below -200.  omega (1/fs) and phi (rad) per component, as the reference
    holds them after reading (read_shape converts its input)."""

    def __init__(self, t_start, t_end, I0, omega, phi, t0, tau):
        self.t0, self.tau = float(t0), float(tau)
        self.t_start = max(float(t_start), self.t0 - 10.0 * self.tau)
        self.t_end = min(float(t_end), self.t0 + 10.0 * self.tau)
        self.I0 = np.asarray(I0, dtype=np.float64)
        self.omega = np.asarray(omega, dtype=np.float64)
        self.phi = np.asarray(phi, dtype=np.float64)

    def __call__(self, t):
        if not (self.t_start <= t < self.t_end):
            return np.zeros(3), 0.0
        x = (t - self.t0) / self.tau
        e = -x * x
        if not e > -200.0:
            return np.zeros(3), 0.0
        return self.I0 * np.cos(self.omega * t + self.phi) * np.exp(e), 0.0


class Algebraic:
    """exc_t_algebraic.f90: I0 / (t - t0)^k inside [t_start, t_end)."""

    def __init__(self, t_start, t_end, I0, t0, k):
        self.t_start, self.t_end = float(t_start), float(t_end)
        self.I0 = np.asarray(I0, dtype=np.float64)
        self.t0, self.k = float(t0), float(k)

    def __call__(self, t):
        if not (self.t_start <= t < self.t_end):
            return np.zeros(3), 0.0
        return self.I0 / (t - self.t0) ** self.k, 0.0


def read_shape(line):
    """A time shape from the reference's input line `name shape t_start t_end
    dim_mode reals...` (the *_read_string routines of exc_t_*.f90), with
    their conversions: EMwave takes a period (fs) and a phase in units of pi
    (exc_t_EMwave.f90:35-36); gauss and EMwave clip the window to t0 -+ 10 tau
    (exc_t_gauss.f90:32-33, exc_t_EMwave.f90:37-38).  Three components."""
    tok = line.split()
    shape, t_start, t_end, dim = tok[1].lower(), float(tok[2]), float(tok[3]), int(tok[4])
    if dim != 3:
        raise ValueError("time shapes of the B order parameter have 3 components")
    r = np.array([float(x) for x in tok[5:]], dtype=np.float64)
    need = {"gauss": dim + 2, "rampe": 2 * dim, "emwave": 3 * dim + 2, "algebraic": dim + 2}
    if shape not in need:
        raise ValueError("unknown time shape %r" % tok[1])
    if len(r) != need[shape]:
        raise ValueError("%s takes %d reals after dim_mode, got %d" % (tok[1], need[shape], len(r)))
    if shape == "gauss":
        return Gauss(t_start, t_end, r[:dim], r[dim], r[dim + 1])
    if shape == "rampe":
        return Rampe(t_start, t_end, r[:dim], r[dim:])
    if shape == "emwave":
        return EMwave(t_start, t_end, r[:dim], 2.0 * PI / r[dim:2 * dim], r[2 * dim:3 * dim] * PI,
                      r[3 * dim], r[3 * dim + 1])
    # algebraic: t0 = real_var(dim+1), k = real_var(dim+2) (exc_t_algebraic.f90:47-49)
    return Algebraic(t_start, t_end, r[:dim], r[dim], r[dim + 1])


class TPulse:
    """Temperature pulse of TPulse.f90 (get_parameter_TPulse defaults, forms
    'exponential' and the cooling form), kT(t) in eV; t_cut_pulse defaults to
    (t_exp - t_frac) / I_over_x / expo + t_exp before the amplitudes are scaled
    by alpha, as there.  The reference defines it without calling it; here it
    feeds a per-step thermal amplitude (llg.spindynamics temperature=)."""

    def __init__(self, T0=150.0, t_start=25, alpha=1.0, I_exp=137.448, I_over_x=286.415, t_exp=22.57, t_frac=19.62,
                 expo=0.47, sigma=None, t_cut_pulse=None, forme="exponential"):
        self.T0, self.t_start, self.t_exp, self.t_frac, self.expo = float(T0), int(t_start), float(t_exp), \
            float(t_frac), float(expo)
        self.sigma = np.sqrt(52.29) if sigma is None else float(sigma)
        self.t_cut = (self.t_exp - self.t_frac) / float(I_over_x) / self.expo + self.t_exp if t_cut_pulse is None \
            else float(t_cut_pulse)
        self.I_exp = float(I_exp) * float(alpha)
        self.I_over_x = float(I_over_x) * float(alpha)
        self.forme = forme

    def temperature(self, t):
        if self.forme == "exponential":
            if t <= self.t_exp - 3.0 * self.sigma or t >= self.t_exp + 3.0 * self.sigma:
                return self.T0
            return self.T0 + self.I_exp * np.exp(-(t - self.t_exp) ** 2 / 2.0 / self.sigma ** 2)
        ts = self.t_start
        if t <= self.t_exp + ts - 3.0 * self.sigma:
            return self.T0
        if t < self.t_exp + ts:
            return self.T0 + self.I_exp * np.exp(-(t - ts - self.t_exp) ** 2 / self.sigma ** 2)
        if t < self.t_cut + ts:
            return self.T0 + self.I_over_x / (t - ts - self.t_frac) ** self.expo
        return self.T0

    def kT(self, t):
        from .lattice import K_B
        return self.temperature(t) * K_B


C_LIGHT = 299.792458          # nm/fs (constant.f90:23)


def laguerre(n, a, x):
    """Generalised Laguerre polynomial L_n^(a)(x) by the three-term recurrence
    (n+1) L_(n+1) = (2n + 1 + a - x) L_n - (n + a) L_(n-1), L_0 = 1,
    L_1 = 1 + a - x (Abramowitz and Stegun 22.7.12).  The reference takes it
    from GSL's gsl_sf_laguerre_n (analytical_function/laguerre.c:12), which is
    absent here."""
    x = np.asarray(x, dtype=np.float64)
    if n == 0:
        return np.ones_like(x)
    l0, l1 = np.ones_like(x), 1.0 + a - x
    for k in range(1, n):
        l0, l1 = l1, ((2 * k + 1 + a - x) * l1 - (k + a) * l0) / (k + 1)
    return l1


class ShapeR:
    """A real-space shape of exc_r.f90: per site an amplitude A and a phase psi
    (the complex factor A exp(i psi) of shape_r, exc_r.f90:27-33) and, for
    the `file` shape, an offset per site.  Shapes: plane, square (|R - c|_inf
    < cutoff), cylinder (|R - c| < cutoff, a sphere in the reference too),
    gaussian (exp(-(|R - c| / cutoff)^2), exponent bounded below by -200),
    laguerre-gauss and its magnetic field components laguerre-gauss-Bx/-By
    (exc_r.f90:196-486), file (offsets from <name>_file, excitation_from_file.f90)."""

    NREAL = {"plane": 0, "square": 4, "cylinder": 4, "gaussian": 4, "laguerre-gauss": 7,
             "laguerre-gauss-bx": 7, "laguerre-gauss-by": 7, "file": 0}

    def __init__(self, line, directory="."):
        tok = line.split()
        self.name, self.shape = tok[0], tok[1]
        key = self.shape.lower()
        if key not in self.NREAL:
            raise ValueError("shape_r identifier not implemented: %r" % self.shape)
        vals = tok[2:2 + self.NREAL[key]]
        if len(vals) != self.NREAL[key]:
            raise ValueError("%s takes %d reals, got %d" % (self.shape, self.NREAL[key], len(vals)))
        self.key = key
        self.offset = None
        if key in ("square", "cylinder", "gaussian"):
            self.center = np.array([float(v) for v in vals[:3]])
            self.cutoff = float(vals[3])
        elif key.startswith("laguerre"):
            self.center = np.array([float(v) for v in vals[:3]])
            self.l, self.p = int(float(vals[3])), int(float(vals[4]))
            self.width_0 = float(vals[5])
            self.wavelength = C_LIGHT * float(vals[6])      # period (fs) -> wavelength (nm)
        elif key == "file":
            path = os.path.join(directory, "%s_%s" % (self.name, self.shape))
            self.offset = np.loadtxt(path, dtype=np.float64).reshape(-1, 3)

    def amplitude_phase(self, R):
        """(A, psi) at the site positions R (N, 3), nm."""
        R = np.asarray(R, dtype=np.float64).reshape(-1, 3)
        n = len(R)
        zero = np.zeros(n)
        if self.key == "plane":
            return np.ones(n), zero
        if self.key == "file":
            return zero, zero
        d = R - self.center
        if self.key == "square":
            return np.all(np.abs(d) < self.cutoff, axis=1).astype(np.float64), zero
        if self.key == "cylinder":
            return (np.linalg.norm(d, axis=1) < self.cutoff).astype(np.float64), zero
        if self.key == "gaussian":
            t = np.linalg.norm(d, axis=1) / self.cutoff
            return np.exp(np.maximum(-t * t, -200.0)), zero
        if self.key == "laguerre-gauss":
            return self._lg(d)
        v = self._lg_bvec(d)
        k = 0 if self.key == "laguerre-gauss-bx" else 2
        return v[:, k] * 1.0e3, v[:, k + 1]                 # V fs / nm^2 -> T

    def _beam(self, d):
        l, p, al = self.l, self.p, abs(self.l)
        z = d[:, 2]
        radius = np.hypot(d[:, 0], d[:, 1])
        phi = np.arctan2(d[:, 1], d[:, 0])
        z_R = PI * self.width_0 ** 2 / self.wavelength
        w_z = self.width_0 * np.sqrt(1.0 + (z / z_R) ** 2)
        with np.errstate(divide="ignore", invalid="ignore"):
            R_curv = np.where(z != 0.0, z * (1.0 + (z_R / np.where(z != 0.0, z, 1.0)) ** 2), np.inf)
        gouy = np.arctan(z / z_R)
        norm = np.sqrt(2.0 ** (al + 1.0) * math.gamma(p + 1.0) / (PI * math.gamma(p + al + 1.0)))
        phase = 2.0 * PI * z / self.wavelength + l * phi - (2 * p + al + 1) * gouy
        phase = np.where(z != 0.0, phase + PI * radius ** 2 / (self.wavelength * R_curv), phase)
        return l, p, al, z, radius, phi, w_z, R_curv, norm, phase

    def _lg(self, d):
        l, p, al, z, r, phi, w_z, R_curv, norm, phase = self._beam(d)
        lag = laguerre(p, al, 2.0 * r ** 2 / w_z ** 2)
        mag = norm / w_z * (r * np.sqrt(2.0) / w_z) ** al * lag * np.exp(-r ** 2 / w_z ** 2)
        far = r ** 2 / w_z ** 2 >= 200.0
        return np.where(far, 0.0, mag), np.where(far, 0.0, phase)

    def _lg_bvec(self, d):
        """(magn_x, phase_x, magn_y, phase_y) of the beam's B (shape_r_laguerre_gauss_Bvec)."""
        l, p, al, z, r, phi, w_z, R_curv, norm, phase = self._beam(d)
        omega = 2.0 * PI * C_LIGHT / self.wavelength
        x = 2.0 * r ** 2 / w_z ** 2
        lag = laguerre(p, al, x)
        lag2 = laguerre(p - 1, al + 1, x) if p >= 1 else np.zeros_like(x)
        e = np.exp(-r ** 2 / w_z ** 2)
        s2 = np.sqrt(2.0) ** al
        with np.errstate(divide="ignore", invalid="ignore"):
            rm1 = r ** (al - 1) if al >= 1 else np.zeros_like(r)
            b_r = 2 * l * norm / omega * rm1 * s2 / w_z ** (al + 1) * lag * e if l != 0 else np.zeros_like(r)
            s = 4.0 / omega * norm * r ** (al + 1) * s2 / w_z ** (al + 3) * lag * e
            if l != 0:
                s = s + -2.0 * al / omega * norm * rm1 * s2 / w_z ** (al + 1) * lag * e
            if p != 0:
                s = s + 8.0 / omega * norm * r ** (al + 1) * s2 / w_z ** (al + 3) * lag2 * e
            c = np.where(z != 0.0, -2.0 / (R_curv * C_LIGHT) * norm * r ** (al + 1) * s2 / w_z ** (al + 1) * lag * e,
                         0.0)
        b_phi = np.sqrt(s ** 2 + c ** 2)
        ph_phi = phase - np.arctan2(s, c)
        # (B_r, B_phi) as complex amplitudes -> (B_x, B_y) (convert_cyl2cart)
        x1 = b_r * np.cos(phase) * np.cos(phi) - b_phi * np.cos(ph_phi) * np.sin(phi)
        x2 = b_r * np.sin(phase) * np.cos(phi) - b_phi * np.sin(ph_phi) * np.sin(phi)
        y1 = b_r * np.cos(phase) * np.sin(phi) + b_phi * np.cos(ph_phi) * np.cos(phi)
        y2 = b_r * np.sin(phase) * np.sin(phi) + b_phi * np.sin(ph_phi) * np.cos(phi)
        out = np.stack([np.hypot(x1, x2), np.arctan2(x2, x1), np.hypot(y1, y2), np.arctan2(y2, y1)], axis=1)
        out[r ** 2 / w_z ** 2 >= 200.0] = 0.0
        return out


def shape_maps(shape_t, shape_r, positions, times):
    """Maps and weights for mj_steps_driven of one excitation: time shape
    shape_t (a read_shape result), real-space shape shape_r (ShapeR) at the
    site positions (Lattice.cell_positions()), evaluated at times (nsteps,
    stages): b_i(t) = T(t) (A_i cos(phi(t) + psi_i) + o_i) as update_exc forms
    it (excitations.f90:164-215)."""
    A, psi = shape_r.amplitude_phase(positions)
    off = shape_r.offset if shape_r.offset is not None else np.zeros((len(A), 3))
    if len(off) != len(A):
        raise ValueError("the file shape holds %d sites, the lattice %d" % (len(off), len(A)))
    return map_weights(shape_t, A, psi, off, times)


def uniform_weights(shape, times):
    """Spatially uniform excitation (A = 1, psi = 0, o = 0): b(t) = T(t) cos phi(t);
    weights for mj_steps_driven with no maps, (nsteps, stages, 1, 3)."""
    w = np.zeros(times.shape + (1, 3))
    for idx in np.ndindex(times.shape):
        T, phi = shape(times[idx])
        w[idx + (0,)] = T * np.cos(phi)
    return w


def map_weights(shape, amplitude, psi, offset, times):
    """One excitation with a real-space shape: maps (3, N, 3) = A cos psi, A sin psi,
    o, and weights (nsteps, stages, 3, 3) = T cos phi, -T sin phi, T."""
    a = np.asarray(amplitude, dtype=np.float64).reshape(-1)
    p = np.asarray(psi, dtype=np.float64).reshape(-1)
    o = np.asarray(offset, dtype=np.float64).reshape(-1, 3)
    maps = np.stack([np.repeat((a * np.cos(p))[:, None], 3, axis=1),
                     np.repeat((a * np.sin(p))[:, None], 3, axis=1), o])
    w = np.zeros(times.shape + (3, 3))
    for idx in np.ndindex(times.shape):
        T, phi = shape(times[idx])
        w[idx + (0,)] = T * np.cos(phi)
        w[idx + (1,)] = -T * np.sin(phi)
        w[idx + (2,)] = T
    return maps, w


def field_of(maps, weights):
    """b_i for each stage: sum_k w_k (x) F_k,i accumulated in map order (the
    device's order); weights (..., k, 3), maps (k, N, 3) -> (..., N, 3)."""
    maps = np.asarray(maps, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)
    out = np.zeros(w.shape[:-2] + maps.shape[1:])
    for k in range(maps.shape[0]):
        out = out + w[..., k, None, :] * maps[k]
    return out
