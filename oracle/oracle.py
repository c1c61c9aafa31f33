"""Python side of the oracle (test infrastructure only).

Wraps liboracle.so (mjoracle.c, the C restatement of the reference's LLG hot
path) with numpy, and holds the numpy restatements that need an FFT:
  dip_conv         fft_H%get_H: pad (fft_ham_internal.f90:26-75), r2c, K^T M per k
                   (fft_ham_fftw.f90:239-262), c2r, crop (fft_ham_internal.f90:77-124);
                   K_F = FFT(K)/prod(N_rep) (fft_ham_fftw.f90:176-186)
  dip_conv_direct  the same circular convolution summed directly (tiny grids)
  dip_dense_field  the reference's dense dipolar term (dipolar_magnetic.f90:181-273)
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
import ctypes
import json
import os
import subprocess

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
HBAR = 6.58211928e-1
K_B = 0.000086173324

_lib = None


def build():
    """Compile liboracle.so (gcc, no FP contraction, OpenMP)."""
    src = os.path.join(HERE, "mjoracle.c")
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
                           "-shared", src, "-o", LIB, "-lm"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER
        L.or_field.argtypes = [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_double)]
        L.or_energy.argtypes = [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_double)]
        L.or_energy.restype = ctypes.c_double
        L.or_steps.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(ctypes.c_double), ctypes.c_int,
                               P(ctypes.c_double)]
        L.or_steps.restype = ctypes.c_int
        L.or_philox4x32_10.argtypes = [P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]
        L.or_normals3.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, P(ctypes.c_double)]
        L.or_thermal_field.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_double,
                                       P(ctypes.c_double)]
        L.or_set_num_threads.argtypes = [ctypes.c_int]
        L.or_div_rn_mismatches.argtypes = [ctypes.c_long, ctypes.c_uint64]
        L.or_div_rn_mismatches.restype = ctypes.c_long
        L.or_num_threads.restype = ctypes.c_int
        L.or_energy_single.argtypes = [ctypes.c_void_p, P(ctypes.c_double), ctypes.c_int, P(ctypes.c_double)]
        L.or_energy_single.restype = ctypes.c_double
        L.or_mc_colouring.argtypes = [ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int),
                                      P(ctypes.c_int)]
        L.or_mc_colouring.restype = ctypes.c_int
        L.or_mc_sweeps.argtypes = [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_int), P(ctypes.c_int),
                                   ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64,
                                   P(ctypes.c_double), ctypes.c_uint32]
        # checks run single-threaded (deterministic, no oversubscription);
        # the timed CPU baseline raises this explicitly
        L.or_set_num_threads(int(os.environ.get("MJ_ORACLE_THREADS", "1")))
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class _Ham(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int),
                ("rp", ctypes.c_void_p), ("ci", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("rpt", ctypes.c_void_p), ("cit", ctypes.c_void_p), ("vt", ctypes.c_void_p),
                ("z", ctypes.c_void_p), ("b", ctypes.c_void_p), ("hdip", ctypes.c_void_p)]


DIPCB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double))


class _Dyn(ctypes.Structure):
    _fields_ = [("integrator", ctypes.c_int), ("timestep", ctypes.c_double), ("damping", ctypes.c_double),
                ("kT", ctypes.c_double), ("seed", ctypes.c_uint64), ("step0", ctypes.c_uint64),
                ("noise", ctypes.c_void_p), ("trunc_eps", ctypes.c_double),
                ("dip", DIPCB), ("dip_user", ctypes.c_void_p),
                ("discretization", ctypes.c_double), ("timesteps", ctypes.c_void_p),
                ("b_stage", ctypes.c_void_p), ("b_end", ctypes.c_void_p), ("rk_order", ctypes.c_int)]


def csr_pair(rows1, cols1, vals, n3):
    """H and H^T as CSR (0-based, sorted columns, duplicates summed)."""
    H = sp.coo_matrix((vals, (rows1 - 1, cols1 - 1)), shape=(n3, n3)).tocsr()
    H.sum_duplicates()
    H.sort_indices()
    T = H.T.tocsr()
    T.sort_indices()
    return H, T


class Oracle:
    """CPU reference model for one Hamiltonian (matjes_amd.lattice.Hamiltonian)."""

    def __init__(self, H, bmap=None):
        self.Hobj = H
        lat = H.lat
        self.n = lat.ncell
        n3 = 3 * self.n
        r, c, v = H.coo_mm()
        self.H, self.HT = csr_pair(r, c, v, n3)
        self._keep = []
        arrs = {}
        for name, M in (("", self.H), ("t", self.HT)):
            arrs["rp" + name] = np.ascontiguousarray(M.indptr.astype(np.int32))
            arrs["ci" + name] = np.ascontiguousarray(M.indices.astype(np.int32))
            arrs["v" + name] = np.ascontiguousarray(M.data.astype(np.float64))
        self.z = None if H.zeeman is None else np.ascontiguousarray(H.zeeman.astype(np.float64))
        self.b = np.ascontiguousarray(H.b_field_sites().reshape(-1))
        self.arrs = arrs
        self.dip = None
        if H.dipolar is not None:
            self.dip = DipolarConv(H.dipolar, lat.dim_lat)
        self.hs = _Ham(self.n,
                       arrs["rp"].ctypes.data, arrs["ci"].ctypes.data, arrs["v"].ctypes.data,
                       arrs["rpt"].ctypes.data, arrs["cit"].ctypes.data, arrs["vt"].ctypes.data,
                       None if self.z is None else self.z.ctypes.data, self.b.ctypes.data, None)

    def _with_dip(self, m):
        hs = _Ham.from_buffer_copy(self.hs)
        hd = None
        if self.dip is not None:
            hd = np.ascontiguousarray(self.dip.h(m.reshape(-1, 3)).reshape(-1))
            hs.hdip = hd.ctypes.data
        return hs, hd

    def field(self, m):
        m = np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(-1))
        B = np.empty_like(m)
        w = np.empty_like(m)
        hs, hd = self._with_dip(m)
        lib().or_field(ctypes.byref(hs), _dp(m), _dp(B), _dp(w))
        return B.reshape(-1, 3)

    def energy(self, m):
        m = np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(-1))
        w = np.empty_like(m)
        hs, hd = self._with_dip(m)
        return lib().or_energy(ctypes.byref(hs), _dp(m), _dp(w))

    def energy_terms(self, m):
        """Energy of each term, as energy_resolved (hamiltonian_collection.f90:365-388):
        every M-M term m.(H_t m) (eval_all, ham_base.f90:181-207), the Zeeman term
        b.(Z m), the FFT term m.(H_fft m) (fft get_E).  Ordered as
        lattice.Hamiltonian.term_names()."""
        H = self.Hobj
        m = np.asarray(m, dtype=np.float64).reshape(-1)
        n3 = 3 * self.n
        out = []
        for blocks in H.mm_terms.values():
            r, c, v = H.coo_mm(blocks=blocks)
            Ht = sp.csr_matrix((v, (r - 1, c - 1)), shape=(n3, n3))
            out.append(float(np.dot(m, Ht @ m)))
        if H.zeeman is not None:
            mm = m.reshape(-1, 3)
            b = self.b.reshape(-1, 3)
            out.append(float(np.sum(b * (mm @ H.zeeman.T))))
        if self.dip is not None:
            out.append(float(np.dot(m, self.dip.h(m.reshape(-1, 3)).reshape(-1))))
        return np.array(out)

    # ---- Metropolis Monte Carlo ----
    def energy_single(self, m, sites1, spins=None):
        """energy_single of each (1-based) site, with its spin or the given ones."""
        m = np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(-1))
        sites1 = np.atleast_1d(sites1)
        out = np.empty(len(sites1))
        for k, i in enumerate(sites1):
            s = m[3 * (int(i) - 1):3 * int(i)] if spins is None else np.asarray(spins, dtype=np.float64).reshape(-1, 3)[k]
            s = np.ascontiguousarray(s, dtype=np.float64)
            out[k] = lib().or_energy_single(ctypes.byref(self.hs), _dp(m), int(i) - 1, _dp(s))
        return out

    def mc_colouring(self):
        lat = self.Hobj.lat
        offs = [o for o in self.Hobj.mm_blocks] + [tuple(-x for x in o) for o in self.Hobj.mm_blocks]
        o3 = np.ascontiguousarray(np.array(offs, dtype=np.int32).reshape(-1))
        dims = np.array(lat.dim_lat, dtype=np.int32)
        per = np.array([int(p) for p in lat.periodic], dtype=np.int32)
        out = np.zeros(3, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        rc = lib().or_mc_colouring(len(offs), o3.ctypes.data_as(ip), dims.ctypes.data_as(ip),
                                   per.ctypes.data_as(ip), out.ctypes.data_as(ip))
        if rc != 0:
            raise ValueError("no colouring")
        return tuple(int(x) for x in out)

    def mc_sweeps(self, m, nsweeps, sampler, kT, seed=0, sweep0=0, state=None, measure_every=0, replica=0):
        """Metropolis sweeps by colour class; state = [cone, cos cone, accepted, proposed, rate]
        (updated in place).  Returns (m, E_total samples, M samples)."""
        if self.dip is not None:
            raise ValueError("no Monte Carlo with the FFT term")
        m = np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(-1)).copy()
        if state is None:
            state = np.array([0.0, 1.0, 0.0, 0.0, 0.0])
        col = np.array(self.mc_colouring(), dtype=np.int32)
        dims = np.array(self.Hobj.lat.dim_lat, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        E, M = [], []
        chunks = [measure_every] * (nsweeps // measure_every) if measure_every > 0 else [nsweeps]
        rest = nsweeps - sum(chunks)
        if rest > 0:
            chunks.append(rest)
        done = 0
        for k, ch in enumerate(chunks):
            lib().or_mc_sweeps(ctypes.byref(self.hs), _dp(m), dims.ctypes.data_as(ip), col.ctypes.data_as(ip),
                               int(ch), int(sampler), float(kT), int(seed), int(sweep0 + done), _dp(state),
                               0x4D430000 + int(replica))
            done += ch
            if measure_every > 0 and ch == measure_every and k < nsweeps // measure_every:
                E.append(self.energy(m))
                M.append(m.reshape(-1, 3).sum(axis=0))
        return m.reshape(-1, 3), np.array(E), np.array(M).reshape(-1, 3)

    def mc_pt(self, m0, ladder, nsweeps, sampler, seed=0, sweep0=0, exchange_every=0, measure_every=0, cone=1.0,
              pt=None):
        """Parallel tempering: replica r swept by or_mc_sweeps with stream 'MC' + r at
        the temperature of its rung; every exchange_every sweeps reorder_temperature
        (paratemp-MC/paratemp.f90:47-121): rung pairs (q, q+1) from q = (n+1) mod 2 for
        the n-th exchange, swapped when delta = (1/kT_{q+1} - 1/kT_q)(E_{q+1} - E_q) > 0
        or u < exp(max(delta, -200)), u from Philox (key seed, counter (q, n, 'PT')).
        pt: the state dict of a previous call to continue from.  Returns
        (pt, E by rung (k, nrep), M by rung (k, nrep, 3)); samples are taken
        before the sweep's exchange."""
        nrep = len(ladder)
        if pt is None:
            pt = {"m": [np.asarray(m0, dtype=np.float64).reshape(-1, 3).copy() for _ in range(nrep)],
                  "st": [np.array([cone, np.cos(cone), 0.0, 0.0, 0.0]) for _ in range(nrep)],
                  "image": list(range(nrep)), "kT": [float(k) for k in ladder], "n_ex": 0,
                  "acc": np.zeros(max(nrep - 1, 1)), "att": np.zeros(max(nrep - 1, 1))}
        Es, Ms = [], []
        two_m32 = 2.0 ** -32
        for s in range(nsweeps):
            for r in range(nrep):
                pt["m"][r], _, _ = self.mc_sweeps(pt["m"][r], 1, sampler, pt["kT"][r], seed=seed, sweep0=sweep0 + s,
                                                  state=pt["st"][r], replica=r)
            ex = exchange_every > 0 and (s + 1) % exchange_every == 0
            me = measure_every > 0 and (s + 1) % measure_every == 0
            if ex or me:
                E = [self.energy(pt["m"][r]) for r in range(nrep)]
            if me:
                Es.append([E[pt["image"][q]] for q in range(nrep)])
                Ms.append([pt["m"][pt["image"][q]].sum(axis=0) for q in range(nrep)])
            if ex and nrep > 1:
                n = pt["n_ex"]
                for q in range((n + 1) % 2, nrep - 1, 2):
                    r1, r2 = pt["image"][q], pt["image"][q + 1]
                    kt1, kt2 = float(ladder[q]), float(ladder[q + 1])
                    delta = (1.0 / kt2 - 1.0 / kt1) * (E[r2] - E[r1])
                    ok = delta > 0.0
                    if not ok:
                        u = philox([q, n & 0xFFFFFFFF, n >> 32, 0x50540000], [seed & 0xFFFFFFFF, seed >> 32])[0]
                        ok = np.exp(max(delta, -200.0)) > u * two_m32
                    pt["att"][q] += 1
                    if ok:
                        pt["image"][q], pt["image"][q + 1] = r2, r1
                        pt["kT"][r1], pt["kT"][r2] = kt2, kt1
                        pt["acc"][q] += 1
                pt["n_ex"] = n + 1
        return pt, np.array(Es).reshape(-1, nrep), np.array(Ms).reshape(-1, nrep, 3)

    def steps(self, m0, nsteps, integrator, timestep, damping, kT=0.0, seed=0, step0=0,
              noise=None, trunc_eps=1.0e-50, energies=False, discretization=0.0, timesteps=None,
              b_stage=None, b_end=None, rk_order=4):
        """timesteps: optional float64 array (nsteps) receiving the timestep of every
        step; discretization > 0 enables update_time (update_time.f90:41-75).
        rk_order: N_order of integration 4 (the reference's Runge-Kutta order N).
        b_stage: optional (nsteps, stages, N, 3) B order parameter of each stage
        (excitations); b_end: optional (nsteps, N, 3) B of the energy after each step."""
        m = np.ascontiguousarray(np.asarray(m0, dtype=np.float64).reshape(-1)).copy()
        E = np.zeros(nsteps) if energies else None
        nz = None
        if noise is not None:
            nz = np.ascontiguousarray(np.asarray(noise, dtype=np.float64).reshape(nsteps, -1))
        cb = DIPCB()
        if self.dip is not None:
            dip = self.dip
            n3 = 3 * self.n

            def _cb(user, mp, out):
                mm = np.ctypeslib.as_array(mp, shape=(n3,)).reshape(-1, 3)
                hh = dip.h(mm).reshape(-1)
                ctypes.memmove(out, hh.ctypes.data, hh.nbytes)
            cb = DIPCB(_cb)
        bs = None if b_stage is None else np.ascontiguousarray(np.asarray(b_stage, dtype=np.float64).reshape(-1))
        be = None if b_end is None else np.ascontiguousarray(np.asarray(b_end, dtype=np.float64).reshape(-1))
        d = _Dyn(int(integrator), float(timestep), float(damping), float(kT), int(seed), int(step0),
                 None if nz is None else nz.ctypes.data, float(trunc_eps), cb, None, float(discretization),
                 None if timesteps is None else timesteps.ctypes.data,
                 None if bs is None else bs.ctypes.data, None if be is None else be.ctypes.data, int(rk_order))
        rc = lib().or_steps(ctypes.byref(self.hs), ctypes.byref(d), _dp(m), int(nsteps),
                            None if E is None else _dp(E))
        if rc != 0:
            raise RuntimeError("or_steps failed %d" % rc)
        return (m.reshape(-1, 3), E) if energies else m.reshape(-1, 3)


# ----------------------------------------------------------------------------
# dipolar (FFT convolution restated with numpy)
# ----------------------------------------------------------------------------
class DipolarConv:
    """H_fft(m) for the operator from matjes_amd.lattice.dipolar_operator."""

    def __init__(self, op, dim_lat):
        self.nrep = tuple(op["N_rep"])
        self.dim = tuple(dim_lat)
        nx, ny, nz = self.nrep
        K = op["K"].reshape(nz, ny, nx, 3, 3)
        nk = float(nx * ny * nz)
        # K_F(a, i, k) = FFT(K(a,i,:))/N
        self.KF = np.fft.rfftn(np.moveaxis(K, (3, 4), (0, 1)), axes=(2, 3, 4)) / nk
        self.Kr = K

    def pad(self, m):
        nx, ny, nz = self.dim
        M = np.zeros((3,) + self.nrep[::-1])
        M[:, :nz, :ny, :nx] = np.moveaxis(m.reshape(nz, ny, nx, 3), 3, 0)
        return M

    def h(self, m):
        M = self.pad(m)
        MF = np.fft.rfftn(M, axes=(1, 2, 3))
        HF = np.einsum("aizyx,azyx->izyx", self.KF, MF)      # H_F(i,k) = sum_a K_F(a,i,k) M_F(a,k)
        Hn = np.fft.irfftn(HF, s=self.nrep[::-1], axes=(1, 2, 3)) * float(np.prod(self.nrep))
        nx, ny, nz = self.dim
        return np.moveaxis(Hn[:, :nz, :ny, :nx], 0, 3).reshape(-1, 3)

    def h_direct(self, m):
        """Circular convolution sum_y K(x-y)^T m(y) done directly (O(N_rep * N))."""
        nx, ny, nz = self.nrep
        M = self.pad(m)
        out = np.zeros((3, nz, ny, nx))
        idx = np.argwhere(np.any(M != 0.0, axis=0))
        for (z, y, x) in idx:
            Kz = np.roll(np.roll(np.roll(self.Kr, z, axis=0), y, axis=1), x, axis=2)
            out += np.einsum("zyxai,a->izyx", Kz, M[:, z, y, x])
        dx, dy, dz = self.dim
        return np.moveaxis(out[:, :dz, :dy, :dx], 0, 3).reshape(-1, 3)


def dip_dense_field(lat, m, period_cutoff=(1, 1, 1)):
    """Field of the reference's dense dipolar term get_dipolar (dipolar_magnetic.f90:181-273):
    B_i = -(H + H^T) m with H(i,j) = -(dip_pref/(8 pi)) mu_i mu_j sum_images (3 r^r^ - 1)/r^3."""
    from matjes_amd.lattice import DIP_PREF, PI
    pos = lat.cell_positions()
    n = len(pos)
    mom = lat.atoms[lat.mag[0]][2]
    a_sc = lat.areal * np.array(lat.dim_lat, dtype=np.float64)[:, None]
    imgs = []
    bnd = [(-period_cutoff[d], period_cutoff[d]) if lat.periodic[d] else (0, 0) for d in range(3)]
    for i3 in range(bnd[2][0], bnd[2][1] + 1):
        for i2 in range(bnd[1][0], bnd[1][1] + 1):
            for i1 in range(bnd[0][0], bnd[0][1] + 1):
                imgs.append(i1 * a_sc[0] + i2 * a_sc[1] + i3 * a_sc[2])
    Hm = np.zeros((n, 3))
    for i in range(n):
        acc = np.zeros(3)
        for img in imgs:
            d = pos - pos[i] + img
            r = np.sqrt(np.sum(d * d, axis=1))
            ok = r > 1e-12
            u = d[ok] / r[ok, None]
            mm = m[ok]
            t = 3.0 * u * np.sum(u * mm, axis=1)[:, None] - mm
            acc += np.sum(t / r[ok, None] ** 3, axis=0)
        Hm[i] = acc * mom * mom * (-DIP_PREF * 0.5 * 0.25 / PI)
    return -2.0 * Hm


# ----------------------------------------------------------------------------
# thermal field and fixtures
# ----------------------------------------------------------------------------
def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def thermal_field(seed, step, n, sigma):
    bt = np.empty(3 * n)
    lib().or_thermal_field(int(seed), int(step), int(n), float(sigma), _dp(bt))
    return bt.reshape(n, 3)


def heun_sigma(setup):
    """sigma of the reference's thermal field for integration 1/2 (spindynamics.f90:266-269)."""
    nl = 1 if setup["integration"] == 1 else 2
    dt = setup["dt"] / float(nl)
    kT = setup.get("T", 0.0) * K_B
    return np.sqrt(setup["alpha"] * kT * HBAR / dt)


def reference_noise(setup, n, first=0, last=None):
    """(sigma, (steps,N,3) thermal fields) of steps first..last-1 of a trajectory
    fixture: heun_sigma times the standard normals of reference_normals."""
    sigma = heun_sigma(setup)
    seed = setup.get("noise_seed", 0)
    last = setup["nsteps"] if last is None else last
    out = np.empty((last - first, n, 3))
    for j in range(first, last):
        out[j - first] = thermal_field(seed, j, n, sigma)
    return sigma, out


def reference_normals(setup, n, step):
    """The standard normals (N,3) that the reference build's driver reads for step
    `step` (0-based) of a trajectory fixture and scales by the sigma it computes
    itself (oracle/ref/mjref_driver.f90); the library's in-kernel generator
    draws the same numbers for key noise_seed and step counter `step`."""
    return thermal_field(setup.get("noise_seed", 0), step, n, 1.0)


def load_fixture(path):
    z = np.load(path, allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["setup"] = json.loads(str(d["setup"]))
    return d
