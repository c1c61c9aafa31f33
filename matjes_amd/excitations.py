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
