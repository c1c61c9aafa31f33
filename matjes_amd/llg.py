"""Host-side mirror of the reference's Hamiltonian and dynamics interfaces.

`Hamiltonian` mirrors `type(hamiltonian)` (src/Hamiltonian_type/
hamiltonian_collection.f90:15-60): terms are registered once, then
`get_eff_field` / `energy` evaluate on the device state.  `spindynamics`
mirrors the step loop of spindynamics_run (src/dynamics/spindynamics.f90:
260-444) with the integrators of `integration` 1 and 2 plus SIB (5) and
Depondt (6).  Every call goes through the C-ABI of libmatjes_hip.so.
"""
import ctypes
import os

import numpy as np

from . import _lib
from .lattice import HBAR, K_B

EULER, HEUN, RK, SIB, DEPONDT = _lib.MJ_EULER, _lib.MJ_HEUN, _lib.MJ_RK, _lib.MJ_SIB, _lib.MJ_DEPONDT


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class Hamiltonian:
    """Device Hamiltonian + state for one lattice (one mj_ctx)."""

    def __init__(self, dim_lat, periodic, device=0, tile=None):
        L = _lib.load()
        self._L = L
        self.dim_lat = tuple(int(x) for x in dim_lat)
        self.periodic = tuple(bool(p) for p in periodic)
        self.ncell = int(np.prod(self.dim_lat))
        ctx = ctypes.c_void_p()
        dims = (ctypes.c_int * 3)(*self.dim_lat)
        per = (ctypes.c_int * 3)(*[int(p) for p in self.periodic])
        _lib.check(L.mj_create(ctypes.byref(ctx), int(device), ctypes.cast(dims, ctypes.c_void_p),
                               ctypes.cast(per, ctypes.c_void_p), 3), "mj_create")
        self._ctx = ctx
        if tile is not None:
            self.set_tile(*tile)

    # ---- lifecycle ----
    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            self._L.mj_destroy(ctypes.byref(self._ctx))
        self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def ctx(self):
        if self._ctx is None or not self._ctx.value:
            raise _lib.MatjesError("context destroyed")
        return self._ctx

    # ---- term registration (set_Hamiltonians / set_fft_Hamiltonians) ----
    def add_coo(self, op_l, op_r, rows1, cols1, vals):
        r = np.ascontiguousarray(rows1, dtype=np.int32)
        c = np.ascontiguousarray(cols1, dtype=np.int32)
        v = _f64(vals)
        _lib.check(self._L.mj_add_hamiltonian_coo(self.ctx, int(op_l), int(op_r), int(len(v)), _ptr(r), _ptr(c),
                                                  _ptr(v)), "mj_add_hamiltonian_coo")

    def add_stencil(self, blocks):
        """blocks: {offset(3): 3x3 block A} with H_MM(i, i+offset) = A."""
        offs = np.ascontiguousarray(np.array(list(blocks.keys()), dtype=np.int32).reshape(-1))
        blk = _f64(np.array([np.asarray(b).reshape(9) for b in blocks.values()]).reshape(-1))
        _lib.check(self._L.mj_add_hamiltonian_stencil(self.ctx, len(blocks), _ptr(offs), _ptr(blk)),
                   "mj_add_hamiltonian_stencil")

    def set_bfield(self, b):
        b = _f64(b).reshape(-1)
        _lib.check(self._L.mj_set_bfield(self.ctx, _ptr(b), len(b)), "mj_set_bfield")

    def set_fft_operator(self, op):
        """op from lattice.dipolar_operator: K (Nk,3,3), N_rep, periodic."""
        K = op["K"]
        kn = np.ascontiguousarray(np.transpose(K, (0, 2, 1)).reshape(-1))   # (a,b,k) column-major
        nrep = (ctypes.c_int * 3)(*op["N_rep"])
        per = (ctypes.c_int * 3)(*[int(p) for p in op["periodic"]])
        _lib.check(self._L.mj_set_fft_operator(self.ctx, 3, ctypes.cast(nrep, ctypes.c_void_p),
                                               ctypes.cast(per, ctypes.c_void_p), _ptr(kn)),
                   "mj_set_fft_operator")

    def load_terms(self, H, form="coo", resolved=False):
        """Register all terms of a lattice.Hamiltonian the way the reference's setup hands them over.

        resolved: one registration per M-M term (exchange, DMI, anisotropy) so that
        energy_terms() reports them apart; otherwise the combined operator at once."""
        if H.mm_blocks:
            parts = list(H.mm_terms.values()) if resolved else [H.mm_blocks]
            for blocks in parts:
                if form == "coo":
                    self.add_coo(_lib.MJ_OP_M, _lib.MJ_OP_M, *H.coo_mm(blocks=blocks))
                else:
                    self.add_stencil(blocks)
        if H.zeeman is not None:
            self.add_coo(_lib.MJ_OP_B, _lib.MJ_OP_M, *H.coo_zeeman())
            self.set_bfield(H.b_field_sites())
        if H.dipolar is not None:
            self.set_fft_operator(H.dipolar)

    def set_tile(self, bx, by, bz):
        _lib.check(self._L.mj_set_tile(self.ctx, int(bx), int(by), int(bz)), "mj_set_tile")

    # ---- state ----
    def set_state(self, m):
        m = _f64(m).reshape(-1)
        _lib.check(self._L.mj_set_state(self.ctx, _ptr(m), len(m)), "mj_set_state")

    def get_state(self):
        m = np.empty(3 * self.ncell)
        _lib.check(self._L.mj_get_state(self.ctx, _ptr(m), len(m)), "mj_get_state")
        return m.reshape(-1, 3)

    def set_state_device(self, ptr):
        _lib.check(self._L.mj_set_state_device(self.ctx, ctypes.c_void_p(ptr), 3 * self.ncell),
                   "mj_set_state_device")

    def get_state_device(self, ptr):
        _lib.check(self._L.mj_get_state_device(self.ctx, ctypes.c_void_p(ptr), 3 * self.ncell),
                   "mj_get_state_device")

    # ---- hamiltonian%get_eff_field / energy ----
    def get_eff_field(self):
        B = np.empty(3 * self.ncell)
        _lib.check(self._L.mj_eff_field(self.ctx, _ptr(B), len(B)), "mj_eff_field")
        return B.reshape(-1, 3)

    def keep_last_field(self, on=True):
        """Have the last stage of every steps() call's last step keep its
        effective field (the field spindynamics.f90:298 evaluates and :400
        prints as Beff_<tag>.dat)."""
        _lib.check(self._L.mj_keep_last_field(self.ctx, 1 if on else 0), "mj_keep_last_field")

    def last_field(self):
        """The kept field of the last stage of the last step, (Ncell, 3)."""
        B = np.empty(3 * self.ncell)
        _lib.check(self._L.mj_last_field(self.ctx, _ptr(B), len(B)), "mj_last_field")
        return B.reshape(-1, 3)

    def energy(self):
        e = np.zeros(1)
        _lib.check(self._L.mj_energy(self.ctx, _ptr(e)), "mj_energy")
        return float(e[0])

    def energy_terms(self):
        """Energy of each registered term (M-M terms in registration order, Zeeman, FFT),
        hamiltonian%energy_resolved (hamiltonian_collection.f90:365-388)."""
        cap = ctypes.c_int(0)
        e = np.zeros(1)
        self._L.mj_energy_terms(self.ctx, _ptr(e), ctypes.byref(cap))   # count only
        e = np.zeros(max(cap.value, 1))
        _lib.check(self._L.mj_energy_terms(self.ctx, _ptr(e), ctypes.byref(cap)), "mj_energy_terms")
        return e[:cap.value].copy()

    def term_mult(self, term, vec, side=0, res=None, alpha=1.0, beta=0.0):
        """mult_r (side 0: alpha H v + beta res) / mult_l (side 1: alpha H^T v + beta res)
        of one registered term (numbered as energy_terms)."""
        v = _f64(np.asarray(vec).reshape(-1))
        out = np.zeros_like(v) if res is None else _f64(np.asarray(res).reshape(-1)).copy()
        _lib.check(self._L.mj_term_mult(self.ctx, int(term), int(side), _ptr(v), len(v), _ptr(out), len(out),
                                        float(alpha), float(beta)), "mj_term_mult")
        return out.reshape(-1, 3)

    # ---- Monte Carlo (Metropolis-MC) ----
    def energy_single(self, sites1, spins=None):
        """hamiltonian%energy_single of each site (1-based indices), with its current
        spin or with the given spins (n,3)."""
        s = np.ascontiguousarray(np.atleast_1d(sites1), dtype=np.int64)
        sp = None if spins is None else _f64(np.asarray(spins).reshape(-1, 3))
        E = np.zeros(len(s))
        _lib.check(self._L.mj_energy_single(self.ctx, len(s), _ptr(s), None if sp is None else _ptr(sp), _ptr(E)),
                   "mj_energy_single")
        return E

    def mc_sweeps(self, nsweeps, sampler, kT, seed=0, sweep0=0, measure_every=0):
        """Metropolis sweeps; returns (E_total, M_sum (k,3)) recorded every measure_every sweeps."""
        ns = nsweeps // measure_every if measure_every > 0 else 0
        E = np.zeros(max(ns, 1))
        M = np.zeros((max(ns, 1), 3))
        _lib.check(self._L.mj_mc_sweeps(self.ctx, int(nsweeps), int(sampler), float(kT), int(seed), int(sweep0),
                                        int(measure_every), _ptr(E), _ptr(M)), "mj_mc_sweeps")
        return E[:ns].copy(), M[:ns].copy()

    def mc_set_cone(self, cone):
        _lib.check(self._L.mj_mc_set_cone(self.ctx, float(cone)), "mj_mc_set_cone")

    def mc_stats(self):
        """{cone, accepted, proposed, rate, colouring (p, a, b)}"""
        out = np.zeros(4)
        col = (ctypes.c_int * 3)()
        _lib.check(self._L.mj_mc_stats(self.ctx, _ptr(out), ctypes.cast(col, ctypes.c_void_p)), "mj_mc_stats")
        return {"cone": out[0], "accepted": out[1], "proposed": out[2], "rate": out[3], "colouring": tuple(col)}

    # ---- parallel tempering Monte Carlo ----
    def mc_pt_setup(self, kT_ladder, cone=1.0):
        kt = _f64(np.asarray(kT_ladder, dtype=np.float64).reshape(-1))
        _lib.check(self._L.mj_mc_pt_setup(self.ctx, len(kt), _ptr(kt), float(cone)), "mj_mc_pt_setup")
        self._nrep = len(kt)

    def mc_pt_sweeps(self, nsweeps, sampler, seed=0, sweep0=0, exchange_every=0, measure_every=0):
        """Returns (E (k, nrep), M (k, nrep, 3)) by rung."""
        nrep = self._nrep
        ns = nsweeps // measure_every if measure_every > 0 else 0
        E = np.zeros(max(ns, 1) * nrep)
        M = np.zeros(max(ns, 1) * nrep * 3)
        _lib.check(self._L.mj_mc_pt_sweeps(self.ctx, int(nsweeps), int(sampler), int(seed), int(sweep0),
                                           int(exchange_every), int(measure_every), _ptr(E), _ptr(M)),
                   "mj_mc_pt_sweeps")
        return E[:ns * nrep].reshape(ns, nrep), M[:ns * nrep * 3].reshape(ns, nrep, 3)

    def mc_pt_state(self):
        nrep = self._nrep
        img = np.zeros(nrep, dtype=np.int32)
        acc = np.zeros(max(nrep - 1, 1))
        att = np.zeros(max(nrep - 1, 1))
        st = np.zeros(4 * nrep)
        _lib.check(self._L.mj_mc_pt_state(self.ctx, _ptr(img), _ptr(acc), _ptr(att), _ptr(st)), "mj_mc_pt_state")
        return {"image_of_rung": img, "swaps_accepted": acc[:nrep - 1], "swaps_attempted": att[:nrep - 1],
                "replicas": st.reshape(nrep, 4)}

    def mc_pt_get_state(self, rung):
        m = np.empty(3 * self.ncell)
        _lib.check(self._L.mj_mc_pt_get_state(self.ctx, int(rung), _ptr(m), len(m)), "mj_mc_pt_get_state")
        return m.reshape(-1, 3)

    # ---- dynamics ----
    def set_rk_order(self, n_order):
        """N_order of integration 4 (the reference's Runge-Kutta order N)."""
        _lib.check(self._L.mj_set_rk_order(self.ctx, int(n_order)), "mj_set_rk_order")

    def steps(self, nsteps, integrator, timestep, damping, kT=0.0, seed=0, step0=0, energies=False):
        E = np.zeros(max(nsteps, 1)) if energies else None
        _lib.check(self._L.mj_steps(self.ctx, int(nsteps), int(integrator), float(timestep), float(damping),
                                    float(kT), int(seed), int(step0), None if E is None else _ptr(E)),
                   "mj_steps")
        return E[:nsteps] if energies else None

    def steps_temperature(self, kT_steps, integrator, timestep, damping, seed=0, step0=0, energies=False):
        """One step per entry of kT_steps (eV), each with its own thermal
        amplitude (mj_steps_temperature)."""
        kts = _f64(np.asarray(kT_steps, dtype=np.float64).reshape(-1))
        n = len(kts)
        E = np.zeros(max(n, 1)) if energies else None
        _lib.check(self._L.mj_steps_temperature(self.ctx, int(n), int(integrator), float(timestep), float(damping),
                                                _ptr(kts), int(seed), int(step0), None if E is None else _ptr(E)),
                   "mj_steps_temperature")
        return E[:n] if energies else None

    def steps_adaptive(self, nsteps, integrator, timestep, damping, discretization, kT=0.0, seed=0, step0=0,
                       energies=False):
        """Steps with update_time (update_time.f90:41-75).  Returns (next timestep,
        timesteps used per step[, energies])."""
        ts = ctypes.c_double(float(timestep))
        used = np.empty(nsteps)
        E = np.empty(nsteps) if energies else None
        _lib.check(self._L.mj_steps_adaptive(self.ctx, int(nsteps), int(integrator), ctypes.byref(ts),
                                             float(damping), float(kT), int(seed), int(step0), float(discretization),
                                             None if E is None else _ptr(E), _ptr(used)), "mj_steps_adaptive")
        return (ts.value, used, E) if energies else (ts.value, used)

    def set_bfield_maps(self, maps):
        """Site maps F_k (k, N, 3) of the time-dependent B order parameter (at most 4)."""
        f = _f64(np.asarray(maps, dtype=np.float64).reshape(-1))
        k = 0 if f.size == 0 else f.size // (3 * self.ncell)
        _lib.check(self._L.mj_set_bfield_maps(self.ctx, int(k), _ptr(f) if f.size else None, f.size),
                   "mj_set_bfield_maps")
        self._nmaps = k

    def steps_driven(self, nsteps, integrator, timestep, damping, weights, kT=0.0, seed=0, step0=0,
                     weights_end=None, energies=False):
        """Steps with b_i = sum_k w_k (x) F_k,i per stage; weights (nsteps, stages, max(k,1), 3)."""
        w = _f64(np.asarray(weights, dtype=np.float64).reshape(-1))
        we = None if weights_end is None else _f64(np.asarray(weights_end, dtype=np.float64).reshape(-1))
        E = np.zeros(max(nsteps, 1)) if energies else None
        _lib.check(self._L.mj_steps_driven(self.ctx, int(nsteps), int(integrator), float(timestep), float(damping),
                                           float(kT), int(seed), int(step0), _ptr(w),
                                           None if we is None else _ptr(we), None if E is None else _ptr(E)),
                   "mj_steps_driven")
        return E[:nsteps] if energies else None

    def steps_noise(self, nsteps, integrator, timestep, damping, kT, noise, energies=False):
        nz = _f64(noise).reshape(-1)
        if nz.size != nsteps * 3 * self.ncell:
            raise ValueError("noise must hold nsteps x 3N values")
        E = np.zeros(max(nsteps, 1)) if energies else None
        _lib.check(self._L.mj_steps_noise(self.ctx, int(nsteps), int(integrator), float(timestep), float(damping),
                                          float(kT), _ptr(nz), None if E is None else _ptr(E)),
                   "mj_steps_noise")
        return E[:nsteps] if energies else None

    def profile_steps(self, nsteps, integrator, timestep, damping, kT=0.0, seed=0, energies=False):
        """(total ms, ms per kernel kind, launches per kind) of nsteps (mj_profile_steps)."""
        tot = np.zeros(1)
        kms = np.zeros(3)
        kc = np.zeros(3, dtype=np.int32)
        _lib.check(self._L.mj_profile_steps(self.ctx, int(nsteps), int(integrator), float(timestep),
                                            float(damping), float(kT), int(seed), 1 if energies else 0, _ptr(tot),
                                            _ptr(kms), _ptr(kc)),
                   "mj_profile_steps")
        return float(tot[0]), kms, kc

    def thermal_normals(self, seed, step):
        out = np.empty(3 * self.ncell)
        _lib.check(self._L.mj_thermal_normals(self.ctx, int(seed), int(step), _ptr(out), len(out)),
                   "mj_thermal_normals")
        return out.reshape(-1, 3)


def from_case(case, device=0, form="coo", tile=None, resolved=False):
    """Device Hamiltonian with the terms and initial state of a workloads.build_case() result."""
    lat = case["lat"]
    h = Hamiltonian(lat.dim_lat, lat.periodic, device=device, tile=tile)
    h.load_terms(case["H"], form=form, resolved=resolved)
    h.set_state(case["m0"])
    return h


def spindynamics(h, nsteps, integrator, timestep, damping, T=0.0, efreq=100, seed=0, step0=0, outdir=None,
                 gra_freq=0, H_ext=(0.0, 0.0, 0.0), gra_beff=False, temperature=None, print_rnd=False):
    """The reference's step loop with its output cadence (spindynamics.f90:260-444).

    Rows are taken after the steps j with mod(j-1, efreq) == 0 (j = 1, 1+efreq,
    ...), labelled with real time (j-1)*timestep, as the reference writes
    EM.dat (spindynamics.f90:336-360).  Returns the records (step, real_time_fs,
    E/Ncell, M, Mx, My, Mz).  With outdir, writes EM.dat there in the
    reference's format and, every gra_freq steps (mod(j-1, gra_freq) == 0),
    magnetic_<j/gra_freq>.dat (write_config, spindynamics.f90:379-409) and,
    with gra_beff (the input key gra_Beff, inp_rw.f90:29), Beff_<j/gra_freq>.dat:
    the effective field of step j's last stage, thermal term excluded
    (print_Beff, spindynamics.f90:298,400).  temperature: an object with
    kT(t) and temperature(t) (excitations.TPulse) evaluated at each step's
    start time (step0 + j - 1) * timestep in place of the constant T.
    print_rnd (the input key print_rnd, random_GNU.f90:146): with gra_freq,
    rnd_num_<j/gra_freq> holds the thermal draws of the last step that drew
    one (kT damping > 1e-10, spindynamics.f90:267-269; zeros before any),
    sigma times the step's Philox normals in the order of the reference's
    list, (3, N) (random_numbers%print_base, spindynamics.f90:424).
    """
    from . import emio
    kT = T * K_B
    beff = bool(outdir and gra_freq and gra_beff)
    rnd = bool(outdir and gra_freq and print_rnd)
    # the thermal amplitude as the device forms it: sqrt(damping kT hbar / dt)
    # with dt = timestep / N_loop for integrations 1 and 2, 2 damping kT hbar / timestep otherwise
    dt_rnd = timestep / 2.0 if integrator == HEUN else timestep
    f_rnd = 1.0 if integrator in (EULER, HEUN) else 2.0
    last_draw = None                       # (Philox step, kT) of the last step that drew
    if beff:
        h.keep_last_field(True)
    out, rows = [], []
    outs = sorted(set([j for j in range(1, nsteps + 1) if (j - 1) % efreq == 0] +
                      ([j for j in range(1, nsteps + 1) if (j - 1) % gra_freq == 0] if gra_freq else [])))
    j = 0
    for target in outs + ([nsteps] if not outs or outs[-1] != nsteps else []):
        k = target - j
        if k <= 0:
            continue
        if temperature is None:
            kts = [kT] * k
            E = h.steps(k, integrator, timestep, damping, kT=kT, seed=seed, step0=step0 + j, energies=True)
        else:
            kts = [temperature.kT((step0 + i) * timestep) for i in range(j, target)]
            E = h.steps_temperature(kts, integrator, timestep, damping, seed=seed, step0=step0 + j, energies=True)
        for i in range(k):
            if kts[i] * damping > 1.0e-10:
                last_draw = (step0 + j + i, kts[i])
        j = target
        if (j - 1) % efreq != 0 and not (gra_freq and (j - 1) % gra_freq == 0):
            continue
        m = h.get_state()
        if (j - 1) % efreq == 0:
            M = m.sum(axis=0) / h.ncell
            out.append((j, (j - 1) * timestep, E[-1], float(np.sqrt(np.sum(M * M))), *M.tolist()))
            Tj = T if temperature is None else temperature.temperature((step0 + j - 1) * timestep)
            rows.append(dict(step=j, real_time=(j - 1) * timestep, e_av=E[-1], M=M, T=Tj, H=H_ext))
        if outdir and gra_freq and (j - 1) % gra_freq == 0:
            emio.write_config(emio.config_name(j // gra_freq, outdir), m)
            if beff:
                emio.write_beff(emio.beff_name(j // gra_freq, outdir), h.last_field())
            if rnd:
                x = np.zeros((h.ncell, 3))
                if last_draw is not None:
                    sigma = np.sqrt(f_rnd * damping * last_draw[1] * HBAR / dt_rnd)
                    x = sigma * h.thermal_normals(seed, last_draw[0])
                emio.write_rnd_num(emio.rnd_num_name(j // gra_freq, outdir), x.reshape(-1))
    if beff:
        h.keep_last_field(False)
    if outdir:
        emio.write_em_dat(os.path.join(outdir, "EM.dat"), rows)
    return out
