"""FP64 reference gradients of an SIB trajectory (test infrastructure only).

The reference has no gradients; what it fixes is the discrete forward map
that the device differentiates: the effective field of the M-M and Zeeman
terms (B = -(H + H^T) m - Z_BM^T b, src/Hamiltonian_type/
hamiltonian_collection.f90:433-454, src/Energy/zeeman.f90:30-86, with the
H_MM of src/Hamiltonian_type/ham_coo.f90:158-200 as restated by
matjes_amd.lattice) and the SIB step (Mentink et al., J. Phys.: Condens.
Matter 22, 176001 (2010); the same step as mjoracle.c's integrator 5).

The map is restated here in numpy so that it runs on complex numbers: every
operation is analytic (no norms, conjugates or branches), so the complex-step
derivative Im L(x + i h v) / h with h = 1e-30 is the exact directional
derivative to rounding, with no cancellation.  That is the reference gradient
the device adjoint is checked against (scalar parameters one by one, per-site
fields and the initial state through random directions).  The real part of
the map is pinned against mjoracle.c's SIB, and the derivative against
central differences and a closed form (single spin in a field).
Only tests/ and bench.py's checks use this.
"""
import numpy as np
import scipy.sparse as sp

from matjes_amd import lattice as L

HBAR = 6.58211928e-1


def unit_terms(lat, names):
    """Unit Hamiltonians of the differentiable parameters.

    names: subset of "J1", "J2", "J3" (exchange of shell k, J = 1 eV), "D"
    (interfacial DMI of shell 1, D = 1 eV, mediator "Ir" below the plane),
    "K" (uniaxial anisotropy along z, K = 1 eV).  Returns lattice.Hamiltonian
    objects (the builders of matjes_amd.lattice, reference conventions)."""
    out = []
    for nm in names:
        H = L.Hamiltonian(lat)
        if nm.startswith("J"):
            H.add_exchange(int(nm[1:]), 1.0)
        elif nm == "D":
            H.add_dmi(1, 1.0, "Ir")
        elif nm == "K":
            H.add_anisotropy([0.0, 0.0, 1.0], 1.0)
        else:
            raise ValueError(nm)
        out.append(H)
    return out


def sym_operator(H):
    """(H + H^T) of the M-M term as a 3N x 3N CSR matrix (AoS site-major)."""
    r, c, v = H.coo_mm()
    n3 = 3 * H.lat.ncell
    M = sp.coo_matrix((v, (r - 1, c - 1)), shape=(n3, n3)).tocsr()
    return (M + M.T).tocsr()


def zeeman_block(lat):
    H = L.Hamiltonian(lat)
    H.add_zeeman([0.0, 0.0, 1.0])
    return np.asarray(H.zeeman, dtype=np.float64)


def _cross(a, b):
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                     a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=1)


def _dot(a, b):
    return np.sum(a * b, axis=1, keepdims=True)


def cayley(m, a):
    """(I + [a]x)^-1 (I - [a]x) m (no conjugation: analytic in a and m)."""
    y = m - _cross(a, m)
    return (y - _cross(a, y) + _dot(a, y) * a) / (1.0 + _dot(a, a))


class SIBMap:
    """The discrete map (theta, b_ext, h, m0) -> m_N of N SIB steps at T = 0."""

    def __init__(self, lat, names):
        self.lat = lat
        self.S = [sym_operator(H) for H in unit_terms(lat, names)]
        self.Z = zeeman_block(lat)
        self.n = lat.ncell

    def field(self, theta, bsites, m):
        acc = np.zeros(3 * self.n, dtype=np.result_type(theta, bsites, m))
        mv = m.reshape(-1)
        for t, S in zip(theta, self.S):
            acc = acc + t * (S @ mv)
        return -(acc.reshape(-1, 3) + bsites @ self.Z)

    def run(self, theta, b_ext, h, m0, nsteps, dt, alpha):
        theta = np.asarray(theta)
        bs = np.asarray(b_ext)[None, :] + (0.0 if h is None else np.asarray(h))
        bs = np.broadcast_to(bs, (self.n, 3))
        f = -0.5 * dt / ((1.0 + alpha * alpha) * HBAR)
        m = np.array(m0)
        for _ in range(nsteps):
            B1 = self.field(theta, bs, m)
            a1 = f * (B1 + alpha * _cross(m, B1))
            p = 0.5 * (m + cayley(m, a1))
            B2 = self.field(theta, bs, p)
            a2 = f * (B2 + alpha * _cross(p, B2))
            m = cayley(m, a2)
        return m


def texture_loss(m, target):
    d = m - target
    return np.sum(d * d)


def complex_step(fun, x, v, h=1.0e-30):
    """Directional derivative d/de fun(x + e v) at e = 0 (fun analytic)."""
    return np.imag(fun(np.asarray(x, dtype=np.complex128) + 1j * h * np.asarray(v))) / h


def reference_gradients(smap, theta, b_ext, h, m0, target, nsteps, dt, alpha, directions=(), seed=0):
    """Exact gradients of texture_loss(m_N, target) with respect to theta and
    b_ext (one complex-step run per component) and directional derivatives
    along `directions` = list of (dh, dm0) pairs (per-site field, initial
    state)."""
    theta = np.asarray(theta, dtype=np.float64)
    b_ext = np.asarray(b_ext, dtype=np.float64)
    hh = np.zeros((smap.n, 3)) if h is None else np.asarray(h, dtype=np.float64)

    def L_of(th, bb, hx, mm):
        return texture_loss(smap.run(th, bb, hx, mm, nsteps, dt, alpha), target)

    gth = np.array([complex_step(lambda t: L_of(t, b_ext, hh, m0), theta, np.eye(len(theta))[k])
                    for k in range(len(theta))])
    gb = np.array([complex_step(lambda b: L_of(theta, b, hh, m0), b_ext, np.eye(3)[k]) for k in range(3)])
    dirs = []
    for dh, dm in directions:
        dirs.append(np.imag(L_of(theta.astype(np.complex128), b_ext.astype(np.complex128),
                                 hh + 1j * 1e-30 * dh, m0 + 1j * 1e-30 * dm)) / 1e-30)
    return gth, gb, np.array(dirs)


def neel_skyrmion(nx, ny, radius, centre=None):
    """Neel skyrmion on an nx x ny grid (site spacing 1): polar angle
    theta = pi (1 - r/R) for r < R, 0 outside (core -z, background +z), in-plane
    part along r-hat (phi = atan2(dy, dx))."""
    cx, cy = ((nx - 1) / 2.0, (ny - 1) / 2.0) if centre is None else centre
    y, x = np.meshgrid(np.arange(ny, dtype=np.float64), np.arange(nx, dtype=np.float64), indexing="ij")
    dx, dy = (x - cx).ravel(), (y - cy).ravel()
    r = np.hypot(dx, dy)
    th = np.where(r < radius, np.pi * (1.0 - r / radius), 0.0)
    phi = np.arctan2(dy, dx)
    return np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)], axis=1)
