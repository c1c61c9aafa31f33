"""Host-side lattice and Hamiltonian-term setup (numpy).

This is the host setup that feeds the device path: it builds the same
Hamiltonian entries the reference's Fortran setup builds, in the same units
(eV, T, fs, nm) and sign conventions, and hands them over in the reference's
own COO form (1-based, (3N)x(3N), op_l/op_r order parameters) or as the
dipolar real-space operator K_n(3,3,Nk).

Reference correspondence:
  Lattice geometry        src/IO-input/rw_lattice.f90:29-36, src/libraries_and_variables/type_lattice.f90:137-181
  atom positions          src/IO-input/rw_cell.f90:20 (fractional -> cartesian)
  neighbour shells        src/setup/neighbor_type.f90:78-133 (get_neighbors), :187-300 (get_neigh_distances)
  exchange J              src/Energy/exchange_Heisenberg_J.f90:92-124 (c_H_J = -1)
  DMI                     src/Energy/exchange_Heisenberg_D.f90:90-132, vector :253-330 (c_H_D = -1)
  anisotropy              src/Energy/anisotropy_heisenberg.f90:148-194, 243-290 (c_H_ani = +1)
  Zeeman ("BM")           src/Energy/zeeman.f90:49-86 (c_zeeman = -1)
  COO expansion           src/Hamiltonian_type/ham_coo.f90:158-200, src/Energy/setH_util.f90:10-36
  dipolar FFT operator    src/Energy/dipolar_magnetic.f90:58-177, 275-308; src/Hamiltonian_type/fft_ham_base.f90:191-224
"""
import numpy as np

# src/fixed-parameters/constant.f90:11-28
K_B = 0.000086173324          # eV/K
MU_0 = 1.0
MU_B = 0.0000578838180661     # eV/T
HBAR = 6.58211928e-1          # eV fs
PI = 3.14159265358979323846264338327950288
# src/Energy/dipolar_magnetic.f90:41  (mu_0 mu_B^2 / nm^3 in eV)
DIP_PREF = 6.745817653234234066975e-7


class Lattice:
    """Bravais lattice with a basis; one magnetic moment per magnetic atom.

    net:      3x3, rows are lattice vectors (any length), as the `lattice` input block
    alat:     3 scale factors (nm)
    dim_lat:  (Nx, Ny, Nz)
    periodic: 3 bools (`Periodic_log`)
    atoms:    list of (name, fractional position (3,), moment in mu_B)
    """

    def __init__(self, net, alat, dim_lat, periodic, atoms):
        net = np.array(net, dtype=np.float64)
        a = np.array(alat, dtype=np.float64).copy()
        for j in range(3):                       # rw_lattice.f90:31-36
            ss = np.sqrt(np.sum(net[j] ** 2))
            net[j] = net[j] / ss
            a[j] = a[j] * ss
        self.areal = net * a[:, None]             # type_lattice.f90:156-160 (rows)
        self.dim_lat = tuple(int(x) for x in dim_lat)
        self.periodic = tuple(bool(p) for p in periodic)
        self.ncell = int(np.prod(self.dim_lat))
        self.atoms = []
        for name, frac, mom in atoms:
            pos = np.asarray(frac, dtype=np.float64) @ self.areal   # rw_cell.f90:20
            self.atoms.append((name, pos, float(mom)))
        self.mag = [i for i, at in enumerate(self.atoms) if at[2] != 0.0]
        if len(self.mag) != 1:
            raise ValueError("this setup supports exactly one magnetic atom per cell")
        self.nmag = 1

    # site index (0-based) of cell (ix,iy,iz), x fastest (lat%index_m_1)
    def site(self, ix, iy, iz):
        nx, ny, _ = self.dim_lat
        return ix + nx * (iy + ny * iz)

    def cell_positions(self):
        nx, ny, nz = self.dim_lat
        iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
        return idx @ self.areal + self.atoms[self.mag[0]][1]


def _neigh_offsets(lat, pos1, pos2_list, neighval):
    """Distances/displacements from one atom to a set of basis atoms.

    Follows get_neigh_distances (neighbor_type.f90:187-300): candidate cells
    in [-Ncheck, Ncheck] along each direction that has more than one cell or
    is periodic, distances sorted and grouped with eps = min|a_i| * 1e-7.
    Returns, for each requested distance index n (1 = smallest, 0-distance
    counted), a list of (basis index, (ix,iy,iz)) and the distance.
    """
    ncheck = max(neighval)
    rng = []
    for d in range(3):
        if lat.dim_lat[d] > 1 or lat.periodic[d]:
            rng.append(np.arange(-ncheck, ncheck + 1))
        else:
            rng.append(np.array([0]))
    cands = []
    for i3 in rng[2]:
        for i2 in rng[1]:
            for i1 in rng[0]:
                shift = i1 * lat.areal[0] + i2 * lat.areal[1] + i3 * lat.areal[2]
                for ia, p2 in enumerate(pos2_list):
                    dv = p2 - pos1 + shift
                    cands.append((float(np.sqrt(np.sum(dv * dv))), ia, (int(i1), int(i2), int(i3))))
    eps = float(np.min(np.sqrt(np.sum(lat.areal ** 2, axis=1)))) * 1.0e-7
    cands.sort(key=lambda c: c[0])
    groups, vmax = [], -10.0 * eps
    for c in cands:
        if c[0] > vmax + eps:
            groups.append([])
            vmax = c[0]
        groups[-1].append(c)
    out = []
    for n in neighval:
        if n - 1 >= len(groups):
            raise ValueError("not enough neighbour distances for shell %d" % n)
        g = groups[n - 1]
        out.append(([(c[1], c[2]) for c in g], g[0][0]))
    return out


def shell_offsets(lat, shell):
    """Cell offsets (ix,iy,iz) of magnetic neighbour shell `shell` (1 = nearest)."""
    pm = lat.atoms[lat.mag[0]][1]
    (lst, dist), = _neigh_offsets(lat, pm, [pm], [shell + 1])   # same type: shifted by one
    return [off for _, off in lst], dist


def dmi_vector(lat, off, mediator):
    """Normalised DMI vector of the bond site->site+off (get_DMI, exchange_Heisenberg_D.f90:253-330)."""
    pm = lat.atoms[lat.mag[0]][1]
    p1 = pm.copy()
    p2 = pm + off[0] * lat.areal[0] + off[1] * lat.areal[1] + off[2] * lat.areal[2]
    center = 0.5 * (p1 + p2)
    xs = [lat.atoms[i][1] for i in range(len(lat.atoms)) if lat.atoms[i][0] == mediator]
    if not xs:
        raise ValueError("no mediating atom of type %r" % mediator)
    (lst, _), = _neigh_offsets(lat, center, xs, [1])
    dsum = np.zeros(3)
    for ia, c in lst:
        x = xs[ia] + c[0] * lat.areal[0] + c[1] * lat.areal[1] + c[2] * lat.areal[2]
        d = np.cross(p1 - x, p2 - x)
        nrm = np.sqrt(np.sum(d * d))
        if nrm > 1e-8:
            dsum = dsum + d / nrm
    nrm = np.sqrt(np.sum(dsum * dsum))
    if nrm > 1e-8:
        dsum = dsum / nrm
    return dsum


class Hamiltonian:
    """Terms of the reference's real-space Hamiltonian for one magnetic atom per cell.

    mm_blocks: {offset (ix,iy,iz): 3x3 block A} with H_MM(i, i+offset) = A
               (J, DMI and on-site anisotropy combined, as combine_Hamiltonians does)
    zeeman:    3x3 block Z of the "BM" term, E_Z = sum_i b_i^T Z m_i
    b:         (3,) uniform field in T, or (N,3) per-site field map (the B order parameter)
    """

    # names of the terms as the reference's energy_resolved reports them
    # (the desc of each Hamiltonian term)
    NAME_J = "symmetric magnetic exchange"
    NAME_D = "antisymmetric magnetic exchange"
    NAME_K = "magnetic anisotropy"
    NAME_Z = "Zeeman energy"
    NAME_DIP = "magnetic dipolar"

    def __init__(self, lat):
        self.lat = lat
        self.mm_blocks = {}
        self.mm_terms = {}      # name -> {offset: block}, in the order the terms were added
        self.zeeman = None
        self.b = None
        self.dipolar = None

    def _add(self, off, blk, name):
        off = tuple(int(x) for x in off)
        for d in (self.mm_blocks, self.mm_terms.setdefault(name, {})):
            if off in d:
                d[off] = d[off] + blk
            else:
                d[off] = blk.copy()

    def term_names(self):
        """Names of the terms in the order energy_terms() reports them."""
        out = list(self.mm_terms)
        if self.zeeman is not None:
            out.append(self.NAME_Z)
        if self.dipolar is not None:
            out.append(self.NAME_DIP)
        return out

    def add_exchange(self, shell, J, c_H_J=-1.0):
        offs, _ = shell_offsets(self.lat, shell)
        blk = c_H_J * (J * np.eye(3))
        for off in offs:
            self._add(off, blk, self.NAME_J)
        return offs

    def add_dmi(self, shell, D, mediator, c_H_D=-1.0):
        offs, _ = shell_offsets(self.lat, shell)
        for off in offs:
            v = dmi_vector(self.lat, off, mediator)
            if np.sqrt(np.sum(v * v)) <= 1e-8:
                continue
            h = np.zeros((3, 3))
            h[0, 1] = v[2] * D
            h[2, 0] = v[1] * D
            h[1, 2] = v[0] * D
            h[1, 0] = -v[2] * D
            h[0, 2] = -v[1] * D
            h[2, 1] = -v[0] * D
            self._add(off, c_H_D * h, self.NAME_D)
        return offs

    def add_anisotropy(self, direction, K, c_H_ani=1.0):
        e = np.asarray(direction, dtype=np.float64)
        e = e / np.sqrt(np.sum(e * e))
        blk = np.zeros((3, 3))
        for i1 in range(3):
            for i2 in range(3):
                blk[i1, i2] += K * e[i1] * e[i2]
        self._add((0, 0, 0), blk * c_H_ani, self.NAME_K)

    def add_zeeman(self, b, c_zeeman=-1.0):
        mom = self.lat.atoms[self.lat.mag[0]][2]
        self.zeeman = (MU_0 * MU_B * c_zeeman) * (mom * np.eye(3))
        self.b = np.asarray(b, dtype=np.float64)

    def add_dipolar(self, period_cutoff=(1, 1, 1)):
        self.dipolar = dipolar_operator(self.lat, period_cutoff)

    # -------- exports in the reference's boundary form --------
    def coo_mm(self, drop_rel=1.0e-10, blocks=None):
        """COO (1-based row, col, val) of H_MM over the whole lattice (ham_coo.f90:158-200).

        Per-block entries below drop_rel*max|block| are dropped as get_coo does.
        Bonds that leave an open boundary are omitted (neighbor_type.f90:111-118).
        """
        lat = self.lat
        nx, ny, nz = lat.dim_lat
        iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
        rows, cols, vals = [], [], []
        for off, blk in (self.mm_blocks if blocks is None else blocks).items():
            jx, jy, jz = ix + off[0], iy + off[1], iz + off[2]
            ok = np.ones(ix.shape, dtype=bool)
            for jj, n, p in ((jx, nx, lat.periodic[0]), (jy, ny, lat.periodic[1]), (jz, nz, lat.periodic[2])):
                if not p:
                    ok &= (jj >= 0) & (jj < n)
            si = (ix + nx * (iy + ny * iz))[ok]
            sj = ((jx % nx) + nx * ((jy % ny) + ny * (jz % nz)))[ok]
            thr = np.max(np.abs(blk)) * drop_rel
            for a in range(3):
                for b in range(3):
                    if abs(blk[a, b]) > thr:
                        rows.append(3 * si + a + 1)
                        cols.append(3 * sj + b + 1)
                        vals.append(np.full(si.shape, blk[a, b]))
        if not rows:
            return np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0)
        return (np.concatenate(rows).astype(np.int32), np.concatenate(cols).astype(np.int32),
                np.concatenate(vals).astype(np.float64))

    def coo_zeeman(self):
        n = self.lat.ncell
        r, c, v = [], [], []
        for a in range(3):
            for b in range(3):
                if self.zeeman[a, b] != 0.0:
                    r.append(3 * np.arange(n) + a + 1)
                    c.append(3 * np.arange(n) + b + 1)
                    v.append(np.full(n, self.zeeman[a, b]))
        return (np.concatenate(r).astype(np.int32), np.concatenate(c).astype(np.int32),
                np.concatenate(v))

    def b_field_sites(self):
        """The B order parameter as (N,3)."""
        n = self.lat.ncell
        if self.b is None:
            return np.zeros((n, 3))
        if self.b.ndim == 1:
            return np.tile(self.b, (n, 1))
        return self.b.reshape(n, 3)


def dipolar_operator(lat, period_cutoff=(1, 1, 1)):
    """Real-space dipolar operator K_n(3,3,Nk) and FFT shape (get_dipolar_fft).

    Returns dict(K=(Nk,3,3) with k index x-fastest, N_rep, periodic).
    K(d) = -(dip_pref/(8 pi)) mu_i mu_j (3 r^ r^ - 1)/r^3 summed over the
    supercell images of the periodic directions; the zero-distance term is 0.
    """
    period = tuple(bool(p or n == 1) for p, n in zip(lat.periodic, lat.dim_lat))
    n_rep, kbd = [], []
    for d in range(3):
        n = lat.dim_lat[d]
        if period[d]:
            n_rep.append(n)
            kbd.append((0, n - 1))
        else:
            n_rep.append(2 * n)
            kbd.append((-n + 1, n - 1))
    # supercell images only along lat%periodic (dipolar_magnetic.f90:275-308)
    a_sc = lat.areal * np.array(lat.dim_lat, dtype=np.float64)[:, None]
    imgs = []
    bnd = [(-period_cutoff[d], period_cutoff[d]) if lat.periodic[d] else (0, 0) for d in range(3)]
    for i3 in range(bnd[2][0], bnd[2][1] + 1):
        for i2 in range(bnd[1][0], bnd[1][1] + 1):
            for i1 in range(bnd[0][0], bnd[0][1] + 1):
                imgs.append(i1 * a_sc[0] + i2 * a_sc[1] + i3 * a_sc[2])
    imgs = np.array(imgs)
    nk = int(np.prod(n_rep))
    K = np.zeros((nk, 3, 3))
    r1 = np.arange(kbd[0][0], kbd[0][1] + 1)
    r2 = np.arange(kbd[1][0], kbd[1][1] + 1)
    r3 = np.arange(kbd[2][0], kbd[2][1] + 1)
    I3, I2, I1 = np.meshgrid(r3, r2, r1, indexing="ij")
    I1, I2, I3 = I1.ravel(), I2.ravel(), I3.ravel()
    kidx = (I1 % n_rep[0]) + n_rep[0] * ((I2 % n_rep[1]) + n_rep[1] * (I3 % n_rep[2]))
    base = np.stack([I1, I2, I3], axis=1).astype(np.float64) @ lat.areal
    acc = np.zeros((base.shape[0], 3, 3))
    for img in imgs:
        diff = base + img
        dn = np.sqrt(np.sum(diff * diff, axis=1))
        nz = dn > 0.0
        u = np.zeros_like(diff)
        u[nz] = diff[nz] / dn[nz, None]
        t = 3.0 * u[:, :, None] * u[:, None, :] - np.eye(3)[None]
        inv3 = np.zeros_like(dn)
        inv3[nz] = 1.0 / dn[nz] ** 3
        acc += t * inv3[:, None, None]
    K[kidx] = acc
    mom = lat.atoms[lat.mag[0]][2]
    K = K * (mom * mom)
    K = -K * DIP_PREF * 0.5 * 0.25 / PI
    return {"K": K, "N_rep": tuple(n_rep), "periodic": period}


# ---------------- standard geometries used by the workloads ----------------

def square_film(nx, ny, nz=1, alat=1.0, moment=2.0, mediator_dz=-0.5, periodic=(True, True, False)):
    atoms = [("Fe", (0.0, 0.0, 0.0), moment)]
    if mediator_dz is not None:
        atoms.append(("Ir", (0.0, 0.0, mediator_dz), 0.0))
    return Lattice(np.eye(3), (alat, alat, alat), (nx, ny, nz), periodic, atoms)


def triangular(nx, ny, alat=1.0, moment=2.0, mediator_dz=-0.5):
    net = [[1.0, 0.0, 0.0], [0.5, 0.86602540378443864676, 0.0], [0.0, 0.0, 1.0]]
    atoms = [("Fe", (0.0, 0.0, 0.0), moment), ("Ir", (0.0, 0.0, mediator_dz), 0.0)]
    return Lattice(net, (alat, alat, alat), (nx, ny, 1), (True, True, False), atoms)
