"""Named workloads: lattice + Hamiltonian + initial state from a plain-dict setup.

Energies in eV, fields in T, times in fs, lengths in nm (the reference's units).
Anisotropy K is the reference's input value: E_K = +K (m.e)^2, so an easy
axis is a negative K (src/Energy/anisotropy_heisenberg.f90:243-290).
"""
import numpy as np

from . import lattice as L

# 64x64 deterministic workload Hamiltonian: J=1 meV, interfacial D=0.3 meV,
# easy-axis K=0.05 meV along z, B=0.5 T along z, mu=2 mu_B
SQUARE_JDKB = dict(geometry="square", periodic=[True, True, False], moment=2.0,
                   J=[1.0e-3], D=0.3e-3, mediator_dz=-0.5, K=-0.05e-3, K_axis=[0.0, 0.0, 1.0],
                   B=[0.0, 0.0, 0.5])
# 1024x1024 stochastic workload Hamiltonian: J1..3 = 1, -0.1, -0.05 meV,
# D=0.4 meV, K=0.1 meV easy axis z, B=1 T along z
TRIANGULAR_J3DKB = dict(geometry="triangular", periodic=[True, True, False], moment=2.0,
                        J=[1.0e-3, -0.1e-3, -0.05e-3], D=0.4e-3, mediator_dz=-0.5,
                        K=-0.1e-3, K_axis=[0.0, 0.0, 1.0], B=[0.0, 0.0, 1.0])
# simple-cubic film with dipolar interaction: J=2 meV, K=0.2 meV along z,
# mu=2 mu_B, a=0.4 nm
FILM_DIP = dict(geometry="square", periodic=[True, True, False], moment=2.0, alat=0.4,
                J=[2.0e-3], D=0.0, K=-0.2e-3, K_axis=[0.0, 0.0, 1.0], B=[0.0, 0.0, 0.0],
                dipolar=True, dip_cut=[1, 1, 1])


def random_sphere(n, seed):
    """n unit vectors uniform on S^2 (normalised Gaussian triples)."""
    g = np.random.default_rng(seed).normal(size=(n, 3))
    return g / np.sqrt(np.sum(g * g, axis=1))[:, None]


def tilted_z(n, seed, max_deg=10.0):
    """+z with per-site random tilts up to max_deg, uniform in solid angle."""
    rng = np.random.default_rng(seed)
    cmin = np.cos(np.deg2rad(max_deg))
    ct = rng.uniform(cmin, 1.0, size=n)
    st = np.sqrt(1.0 - ct * ct)
    ph = rng.uniform(0.0, 2.0 * np.pi, size=n)
    return np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)


def make_lattice(s):
    nx, ny, nz = s["size"]
    a = s.get("alat", 1.0)
    dz = s.get("mediator_dz", -0.5) if s.get("D", 0.0) != 0.0 else None
    if s["geometry"] == "triangular":
        if dz is None:
            atoms = [("Fe", (0.0, 0.0, 0.0), s["moment"])]
            net = [[1.0, 0.0, 0.0], [0.5, 0.86602540378443864676, 0.0], [0.0, 0.0, 1.0]]
            return L.Lattice(net, (a, a, a), (nx, ny, nz), s["periodic"], atoms)
        lat = L.triangular(nx, ny, alat=a, moment=s["moment"], mediator_dz=dz)
        if nz != 1 or tuple(s["periodic"]) != (True, True, False):
            lat = L.Lattice([[1.0, 0.0, 0.0], [0.5, 0.86602540378443864676, 0.0], [0.0, 0.0, 1.0]],
                            (a, a, a), (nx, ny, nz), s["periodic"],
                            [("Fe", (0.0, 0.0, 0.0), s["moment"]), ("Ir", (0.0, 0.0, dz), 0.0)])
        return lat
    return L.square_film(nx, ny, nz, alat=a, moment=s["moment"], mediator_dz=dz, periodic=s["periodic"])


def make_hamiltonian(s, lat):
    H = L.Hamiltonian(lat)
    for i, j in enumerate(s.get("J", [])):
        if j != 0.0:
            H.add_exchange(i + 1, j)
    if s.get("D", 0.0) != 0.0:
        H.add_dmi(1, s["D"], "Ir")
    if s.get("K", 0.0) != 0.0:
        H.add_anisotropy(s.get("K_axis", [0.0, 0.0, 1.0]), s["K"])
    B = np.asarray(s.get("B", [0.0, 0.0, 0.0]), dtype=np.float64)
    bmap = None
    if s.get("bmap_sigma"):
        rng = np.random.default_rng(s.get("seed", 0) + 1000003)
        bmap = B[None, :] + s["bmap_sigma"] * rng.normal(size=(lat.ncell, 3))
        H.add_zeeman(bmap)
    elif np.sqrt(np.sum(B * B)) >= 1.0e-8:
        H.add_zeeman(B)
    if s.get("dipolar"):
        H.add_dipolar(tuple(s.get("dip_cut", [1, 1, 1])))
    return H, bmap


def build_case(s):
    lat = make_lattice(s)
    H, bmap = make_hamiltonian(s, lat)
    if s.get("init") == "tilted_z":
        m0 = tilted_z(lat.ncell, s.get("seed", 0), s.get("tilt_deg", 10.0))
    else:
        m0 = random_sphere(lat.ncell, s.get("seed", 0))
    return {"lat": lat, "H": H, "m0": m0, "bmap": bmap}
