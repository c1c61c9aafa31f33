"""HIP path against the reference fixtures and the oracle (MI355X).

Every compute call goes through the C-ABI (include/matjes_hip.h) of
matjes_amd/libmatjes_hip.so.  Tolerances: fields and energies of one state
to 1e-14/1e-13 relative (the stencil sums in a different order than the
reference's matrix backend); trajectories to 1e-12 absolute per component
after hundreds of steps and 1e-10 relative after 10,000 steps.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden

pytestmark = pytest.mark.gpu

FIELD_CASES = ["field_sq8", "field_tri8", "field_sq8_bmap", "field_dip_open"]
TRAJ_CASES = ["traj_sq8_euler", "traj_sq8_heun", "traj_tri8_heun_T10", "traj_tri8_euler_T10", "traj_sq8_heun_bmap",
              "traj_sq8_rk4", "traj_tri8_rk3", "traj_sq8_rk7"]
LONG_CASES = ["traj_sq32_heun_10k", "traj_sq32_euler_10k"]
# finite temperature: the reference scaled the oracle's standard normals by the
# sigma it computed itself (spindynamics.f90:269, oracle/ref/mjref_driver.f90)
THERMAL_CASES = ["traj_tri8_heun_T10", "traj_tri8_euler_T10"]


def _dev(case, form="coo", tile=None, resolved=False):
    from matjes_amd import llg
    return llg.from_case(case, form=form, tile=tile, resolved=resolved)


def _case(s):
    from matjes_amd import workloads
    return workloads.build_case(s)


@pytest.mark.parametrize("form", ["coo", "stencil"])
@pytest.mark.parametrize("name", FIELD_CASES)
def test_field_matches_reference(name, form):
    d = golden(name)
    c = _case(d["setup"])
    with _dev(c, form) as h:
        B = h.get_eff_field().reshape(-1)
        E = h.energy()
    ref = d["field"]
    assert np.max(np.abs(B - ref)) <= 1e-14 * np.max(np.abs(ref))
    assert abs(E - d["energy"]) <= 1e-13 * abs(d["energy"])


@pytest.mark.parametrize("form", ["coo", "stencil"])
@pytest.mark.parametrize("name", FIELD_CASES)
def test_energy_terms_match_reference(name, form):
    """Per-term energies (energy_resolved) against the reference term by term; with the
    terms registered one by one the field and total are still the reference's."""
    from matjes_amd import _lib
    d = golden(name)
    c = _case(d["setup"])
    with _dev(c, form, resolved=True) as h:
        E = h.energy_terms()
        Etot = h.energy()
        B = h.get_eff_field().reshape(-1)
        short = np.zeros(1)
        import ctypes
        cnt = ctypes.c_int(len(E) - 1)
        rc = h._L.mj_energy_terms(h.ctx, short.ctypes.data, ctypes.byref(cnt))
    assert rc == _lib.MJ_EINVAL and cnt.value == len(E)
    names = c["H"].term_names()
    ref = dict(zip((str(t) for t in d["term_names"]), d["term_energies"]))
    assert sorted(names) == sorted(ref)
    scale = np.max(np.abs(d["term_energies"]))
    for nm, e in zip(names, E):
        assert abs(e - ref[nm]) <= 1e-13 * scale, (nm, e, ref[nm])
    assert abs(np.sum(E) - Etot) <= 1e-13 * scale
    assert np.max(np.abs(B - d["field"])) <= 1e-14 * np.max(np.abs(d["field"]))


@pytest.mark.parametrize("size,periodic", [((8, 6, 3), (True, True, False)), ((40, 24, 2), (True, True, False)),
                                           ((20, 18, 1), (False, False, False))])
def test_energy_terms_dipolar_film_vs_oracle(oracle_lib, size, periodic):
    from matjes_amd import workloads
    c = _case(dict(workloads.FILM_DIP, size=list(size), periodic=list(periodic), seed=5))
    with _dev(c, "stencil", resolved=True) as h:
        E = h.energy_terms()
    ref = oracle_lib.Oracle(c["H"]).energy_terms(c["m0"])
    assert len(E) == len(ref) == len(c["H"].term_names())
    assert np.max(np.abs(E - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_energy_terms_generic_path_refused():
    from matjes_amd import _lib, llg, workloads
    c = _case(dict(workloads.SQUARE_JDKB, size=[12, 10, 1], seed=4))
    r, cc, v = c["H"].coo_mm()
    v = v.copy()
    v[::7] *= 1.1
    with llg.Hamiltonian(c["lat"].dim_lat, c["lat"].periodic) as h:
        h.add_coo(1, 1, r, cc, v)
        h.set_state(c["m0"])
        with pytest.raises(_lib.MatjesError):
            h.energy_terms()


@pytest.mark.parametrize("tile", [(64, 4, 1), (32, 8, 1), (16, 16, 1), (8, 8, 1)])
def test_field_tiles(oracle_lib, tile):
    from matjes_amd import workloads
    c = _case(dict(workloads.TRIANGULAR_J3DKB, size=[40, 24, 1], seed=3))
    with _dev(c, "coo", tile=tile) as h:
        B = h.get_eff_field()
    ref = oracle_lib.Oracle(c["H"]).field(c["m0"])
    assert np.max(np.abs(B - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_dipolar_periodic_film_vs_oracle(oracle_lib):
    from matjes_amd import workloads
    c = _case(dict(workloads.FILM_DIP, size=[8, 6, 3], periodic=[True, True, False], seed=8))
    with _dev(c) as h:
        B = h.get_eff_field()
        E = h.energy()
    o = oracle_lib.Oracle(c["H"])
    ref = o.field(c["m0"])
    assert np.max(np.abs(B - ref)) <= 1e-13 * np.max(np.abs(ref))
    assert abs(E - o.energy(c["m0"])) <= 1e-12 * abs(E)


def _traj_gpu(d, c):
    from oracle import oracle as orc
    s = d["setup"]
    kT = s.get("T", 0.0) * orc.K_B
    noise = orc.reference_noise(s, c["lat"].ncell)[1] if s.get("T", 0.0) > 0.0 else None
    out, last = [], 0
    with _dev(c) as h:
        if s["integration"] == 4:
            h.set_rk_order(s["N_order"])
        for st in d["steps"]:
            k = int(st - last)
            if noise is None:
                E = h.steps(k, s["integration"], s["dt"], s["alpha"], kT=kT, energies=True)
            else:
                E = h.steps_noise(k, s["integration"], s["dt"], s["alpha"], kT, noise[last:st], energies=True)
            out.append((h.get_state(), E[-1]))
            last = st
    return out


@pytest.mark.parametrize("name", TRAJ_CASES)
def test_trajectory_matches_reference(name):
    d = golden(name)
    c = _case(d["setup"])
    for k, (m, E) in enumerate(_traj_gpu(d, c)):
        assert np.max(np.abs(m - d["m"][k])) <= 1e-12, k
        assert abs(E - d["E"][k]) <= 1e-12 * abs(d["E"][k]), k


def _traj_gpu_philox(d, c, tile=None):
    """The production thermal path: mj_steps draws the normals in-kernel
    (Philox4x32-10, key noise_seed, step counter = step index) and forms the
    amplitude sqrt(alpha kT hbar / dt_stage) itself."""
    from oracle import oracle as orc
    s = d["setup"]
    kT = s.get("T", 0.0) * orc.K_B
    out, last = [], 0
    with _dev(c, tile=tile) as h:
        for st in d["steps"]:
            E = h.steps(int(st - last), s["integration"], s["dt"], s["alpha"], kT=kT, seed=s["noise_seed"],
                        step0=last, energies=True)
            out.append((h.get_state(), E[-1]))
            last = st
    return out


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6, 7])
def test_runge_kutta_order_n_vs_oracle(oracle_lib, order):
    """Integration 4 on the stage kernels (per-Hamiltonian compiled lattice
    size, a per-site field) against the oracle for every order the reference
    has a Butcher table for; orders 1 and 2 are integrations 1 and 2 bit for
    bit (the same tables, solver_order.f90:65-78)."""
    from matjes_amd import llg, workloads
    c = _case(dict(workloads.SQUARE_JDKB, size=[72, 64, 1], seed=13, bmap_sigma=0.2))
    with _dev(c) as h:
        h.set_rk_order(order)
        E = h.steps(40, llg.RK, 1.0, 0.1, energies=True)
        m = h.get_state()
    mo, Eo = oracle_lib.Oracle(c["H"]).steps(c["m0"], 40, 4, 1.0, 0.1, energies=True, rk_order=order)
    assert np.max(np.abs(m - mo)) <= 1e-12
    assert np.max(np.abs(E - Eo)) <= 1e-12 * np.max(np.abs(Eo))
    if order in (1, 2):
        with _dev(c) as h:
            E2 = h.steps(40, order, 1.0, 0.1, energies=True)
            assert np.array_equal(h.get_state(), m) and np.array_equal(E2, E)


def test_runge_kutta_refuses_a_thermal_field():
    from matjes_amd import _lib, llg, workloads
    c = _case(dict(workloads.SQUARE_JDKB, size=[16, 16, 1], seed=14))
    with _dev(c) as h:
        with pytest.raises(_lib.MatjesError):
            h.steps(2, llg.RK, 1.0, 0.1, kT=10.0 * 8.617e-5)


@pytest.mark.parametrize("name", THERMAL_CASES)
def test_in_kernel_thermal_field_matches_reference(name):
    """In-kernel normals and amplitude against the reference's own sigma: the
    trajectory equals the reference's to 1e-12 (its normals are the oracle's
    libm Box-Muller draws of the same Philox words, a few ulp from the
    device's)."""
    d = golden(name)
    c = _case(d["setup"])
    for k, (m, E) in enumerate(_traj_gpu_philox(d, c)):
        assert np.max(np.abs(m - d["m"][k])) <= 1e-12, k
        assert abs(E - d["E"][k]) <= 1e-12 * abs(d["E"][k]), k


def test_in_kernel_thermal_field_10k_steps_matches_reference():
    """10,000 steps at 10 K of the triangular workload's Hamiltonian and
    integrator (32x32), in-kernel thermal field: 1e-10 per spin relative
    against the reference through 5000 steps, 1e-8 at 7500 and 10,000 (the
    chaotic growth of last-bit differences that the oracle, drawing the
    reference's own normals, shows as well: 1.5e-12, 4.5e-11, 6.3e-10, 1.3e-9);
    energy per site to 1e-12 through 5000 steps and, with the same growth,
    1e-10 at 7500 and 10,000 (measured 1.0e-12 at 7500)."""
    if not os.path.exists(os.path.join(GOLDEN, "traj_tri32_heun_T10_10k.npz")):
        pytest.skip("fixture not generated")
    d = golden("traj_tri32_heun_T10_10k")
    c = _case(d["setup"])
    bound = [1e-10, 1e-10, 1e-8, 1e-8]
    ebound = [1e-12, 1e-12, 1e-10, 1e-10]
    for k, (m, E) in enumerate(_traj_gpu_philox(d, c)):
        rel = np.max(np.linalg.norm(m - d["m"][k], axis=1) / np.linalg.norm(d["m"][k], axis=1))
        assert rel <= bound[k], (k, rel)
        assert abs(E - d["E"][k]) <= ebound[k] * abs(d["E"][k]), (k, E, d["E"][k])


@pytest.mark.parametrize("name", LONG_CASES)
def test_trajectory_10k_steps_matches_reference(name):
    if not os.path.exists(os.path.join(GOLDEN, name + ".npz")):
        pytest.skip("fixture not generated")
    d = golden(name)
    c = _case(d["setup"])
    for k, (m, E) in enumerate(_traj_gpu(d, c)):
        # error per spin relative to the spin length (|m| = 1)
        rel = np.max(np.linalg.norm(m - d["m"][k], axis=1) / np.linalg.norm(d["m"][k], axis=1))
        assert rel <= 1e-10, (k, rel)
        assert abs(E - d["E"][k]) <= 1e-10 * abs(d["E"][k])


@pytest.mark.parametrize("integ", [1, 2, 5, 6])
@pytest.mark.parametrize("T", [0.0, 30.0])
def test_integrators_vs_oracle(oracle_lib, integ, T):
    """Same injected thermal field on both sides; all four integrators."""
    from matjes_amd import workloads
    s = dict(workloads.TRIANGULAR_J3DKB, size=[16, 12, 1], seed=6)
    c = _case(s)
    n, steps, dt, alpha = c["lat"].ncell, 150, 0.5, 0.05
    kT = T * oracle_lib.K_B
    noise = None
    if T > 0:
        noise = np.stack([oracle_lib.thermal_field(11, j, n, 0.01) for j in range(steps)])
    o = oracle_lib.Oracle(c["H"])
    mo, Eo = o.steps(c["m0"], steps, integ, dt, alpha, kT=kT, noise=noise, energies=True)
    with _dev(c) as h:
        if noise is None:
            Eg = h.steps(steps, integ, dt, alpha, kT=kT, energies=True)
        else:
            Eg = h.steps_noise(steps, integ, dt, alpha, kT, noise, energies=True)
        mg = h.get_state()
    assert np.max(np.abs(mg - mo)) <= 1e-12
    assert np.max(np.abs(Eg - Eo)) <= 1e-12 * max(np.max(np.abs(Eo)), 1e-3)


def test_thermal_normals_vs_oracle(oracle_lib):
    """The device's normals (hardware-seeded quotient and square root in the
    Box-Muller transform) against the oracle's libm ones of the same Philox
    words: a few ulp of the largest normal (|g| <= 6.7)."""
    from matjes_amd import workloads
    c = _case(dict(workloads.SQUARE_JDKB, size=[256, 256, 1], seed=1))
    with _dev(c) as h:
        for step in (0, 1, 123456789012):
            g = h.thermal_normals(0xC0FFEE, step)
            want = oracle_lib.thermal_field(0xC0FFEE, step, c["lat"].ncell, 1.0)
            assert np.max(np.abs(g - want)) <= 8e-15


def test_site_algebra_sqrt_and_reciprocal_are_the_library_ones():
    """The stage kernels' sqrt and 1 / x (the compiler's sequences without the
    range scaling) give the same bits as sqrt and 1.0 / x over 8 decades
    around 1 and the field magnitudes, and sqrt(0) = 0."""
    import ctypes
    from matjes_amd import _lib, workloads
    rng = np.random.default_rng(17)
    x = np.concatenate([10.0 ** rng.uniform(-12, 12, 1 << 20), 1.0 + rng.uniform(-1e-6, 1e-6, 1 << 18),
                        [1.0, 2.0, 0.5, np.nextafter(1.0, 2.0), np.nextafter(1.0, 0.0), 1e-8, 1e8]])
    x = np.ascontiguousarray(x)
    c = _case(dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=1))
    with _dev(c) as h:
        def probe(fn, v):
            y = np.empty_like(v)
            _lib.check(_lib.load().mj_math_probe(h.ctx, fn, len(v), ctypes.c_void_p(v.ctypes.data),
                                                 ctypes.c_void_p(y.ctypes.data)), "mj_math_probe")
            return y
        assert np.array_equal(probe(0, x), probe(1, x))
        assert np.array_equal(probe(2, x), probe(3, x))
        assert np.array_equal(probe(0, np.zeros(4)), np.zeros(4))
        # and both equal numpy's correctly rounded results
        assert np.array_equal(probe(0, x), np.sqrt(x))
        assert np.array_equal(probe(2, x), 1.0 / x)


@pytest.mark.parametrize("integ", [2, 5])
def test_philox_noise_trajectory_vs_oracle(oracle_lib, integ):
    """In-kernel Philox thermal field against the oracle's Philox field."""
    from matjes_amd import workloads
    c = _case(dict(workloads.TRIANGULAR_J3DKB, size=[16, 16, 1], seed=2))
    kT = 10.0 * oracle_lib.K_B
    o = oracle_lib.Oracle(c["H"])
    mo = o.steps(c["m0"], 100, integ, 0.5, 0.05, kT=kT, seed=99, step0=5)
    with _dev(c) as h:
        h.steps(100, integ, 0.5, 0.05, kT=kT, seed=99, step0=5)
        mg = h.get_state()
    assert np.max(np.abs(mg - mo)) <= 1e-11


def _last_stage_state(o, m_prev, integ):
    """The state the last stage of a step evaluates its field on: the state
    entering the step (integration 1), or the first stage's result
    (integration 2: Euler from it with half the timestep, no thermal term,
    which the reference adds in the last stage only)."""
    return m_prev if integ == 1 else o.steps(m_prev, 1, 1, 0.25, 0.05)


@pytest.mark.parametrize("integ,T,fuse", [(1, 0.0, "1"), (1, 10.0, "1"), (2, 0.0, "1"), (2, 10.0, "1"),
                                          (2, 10.0, "0")])
def test_last_stage_field_vs_oracle(oracle_lib, integ, T, fuse):
    """mj_keep_last_field / mj_last_field: the effective field of the last
    stage of a call's last step (the Beff that print_Beff writes, thermal term
    excluded) against the oracle's field of that stage's state, after a call
    of 70 steps and after a further one-step call (1440 sites: the
    ahead-of-time kernels; the compiled and fused kernels' field is compared
    with these in test_gpu_jit); 1e-12 relative.  A new state drops the kept
    field."""
    from matjes_amd import _lib, workloads
    c = _case(dict(workloads.TRIANGULAR_J3DKB, size=[40, 36, 1], seed=5))
    kT = T * oracle_lib.K_B
    o = oracle_lib.Oracle(c["H"])
    old = os.environ.get("MJ_FUSE")
    os.environ["MJ_FUSE"] = fuse
    try:
        with _dev(c, "stencil") as h:
            h.keep_last_field(True)
            for n, step0 in ((70, 0), (1, 70)):
                h.steps(n, integ, 0.5, 0.05, kT=kT, seed=9, step0=step0)
                Bg = h.last_field()
                m_prev = o.steps(c["m0"], step0 + n - 1, integ, 0.5, 0.05, kT=kT, seed=9, step0=0)
                Bo = o.field(_last_stage_state(o, m_prev, integ))
                assert np.max(np.abs(Bg - Bo)) <= 1e-12 * np.max(np.abs(Bo)), (n, step0)
            h.set_state(c["m0"])
            with pytest.raises(_lib.MatjesError):
                h.last_field()
    finally:
        if old is None:
            os.environ.pop("MJ_FUSE", None)
        else:
            os.environ["MJ_FUSE"] = old


def test_spindynamics_writes_beff_files(oracle_lib, tmp_path):
    """spindynamics with gra_Beff: Beff_<j/gra_freq>.dat at the steps j with
    mod(j-1, gra_freq) = 0, in print_Beff's format, holding the field of step
    j's last stage (12 significant digits printed)."""
    from matjes_amd import emio, llg, workloads
    c = _case(dict(workloads.TRIANGULAR_J3DKB, size=[16, 12, 1], seed=6))
    o = oracle_lib.Oracle(c["H"])
    with _dev(c, "stencil") as h:
        llg.spindynamics(h, 25, 2, 0.5, 0.05, T=0.0, efreq=5, outdir=str(tmp_path), gra_freq=12, gra_beff=True)
    for tag, j in ((0, 1), (1, 13), (2, 25)):
        got = np.loadtxt(emio.beff_name(tag, str(tmp_path)))
        m_prev = o.steps(c["m0"], j - 1, 2, 0.5, 0.05) if j > 1 else c["m0"]
        want = o.field(_last_stage_state(o, m_prev, 2))
        assert got.shape == want.shape
        assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want)), tag


@pytest.mark.parametrize("integ,pulse", [(2, False), (6, False), (2, True)])
def test_spindynamics_writes_rnd_num_files(oracle_lib, tmp_path, integ, pulse):
    """spindynamics with print_rnd: rnd_num_<j/gra_freq> at the steps j with
    mod(j-1, gra_freq) = 0 holds the last drawn thermal list, (3, N) order, in
    print_base's format: sigma (the reference's amplitude for Heun; 2 alpha kT
    hbar / dt for Depondt) times the oracle's Philox normals of that step; with
    a temperature pulse from 0 K the first file holds zeros (no draw yet)."""
    from matjes_amd import emio, excitations, llg, workloads
    from matjes_amd.lattice import HBAR
    c = _case(dict(workloads.TRIANGULAR_J3DKB, size=[16, 12, 1], seed=6))
    n, alpha, ts = c["lat"].ncell, 0.05, 0.5
    tp = excitations.TPulse(T0=0.0, I_exp=40.0) if pulse else None
    with _dev(c, "stencil") as h:
        llg.spindynamics(h, 25, integ, ts, alpha, T=10.0, efreq=5, seed=9, outdir=str(tmp_path), gra_freq=12,
                         print_rnd=True, temperature=tp)
    dt = ts / 2.0 if integ == 2 else ts
    f = 1.0 if integ == 2 else 2.0
    for tag, j in ((0, 1), (1, 13), (2, 25)):
        got = np.loadtxt(emio.rnd_num_name(tag, str(tmp_path)))
        kT = tp.kT((j - 1) * ts) if pulse else 10.0 * oracle_lib.K_B
        if kT * alpha <= 1e-10:
            assert tag == 0 and pulse and got.shape == (3 * n,) and not np.any(got)
            continue
        want = oracle_lib.thermal_field(9, j - 1, n, np.sqrt(f * alpha * kT * HBAR / dt)).reshape(-1)
        assert got.shape == want.shape
        assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want)), tag


@pytest.mark.parametrize("size", [(96, 80), (130, 70), (97, 70)])
def test_compiled_kernels_tile_mix_vs_oracle(oracle_lib, size):
    """The per-Hamiltonian kernels on lattices with interior, boundary and
    partial tiles: 96x80 and 130x70 fill interior tiles straight from HBM into
    LDS (even rows), 97x70 (odd rows) through registers; Heun with the Philox
    field and the energy every step, and the field after it, against the
    oracle."""
    from matjes_amd import workloads
    c = _case(dict(workloads.TRIANGULAR_J3DKB, size=[size[0], size[1], 1], seed=7))
    kT = 10.0 * oracle_lib.K_B
    o = oracle_lib.Oracle(c["H"])
    mo, Eo = o.steps(c["m0"], 20, 2, 0.5, 0.05, kT=kT, seed=11, step0=3, energies=True)
    with _dev(c) as h:
        E = h.steps(20, 2, 0.5, 0.05, kT=kT, seed=11, step0=3, energies=True)
        mg = h.get_state()
        B = h.get_eff_field()
    assert np.max(np.abs(mg - mo)) <= 1e-11
    assert np.max(np.abs(E - Eo)) <= 1e-11 * np.max(np.abs(Eo))
    ref = o.field(mo)
    assert np.max(np.abs(B - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_generic_sparse_path(oracle_lib):
    """A Hamiltonian that is not translation invariant takes the generic path."""
    from matjes_amd import workloads
    c = _case(dict(workloads.SQUARE_JDKB, size=[12, 10, 1], seed=4))
    H = c["H"]
    r, cc, v = H.coo_mm()
    v = v.copy()
    v[::7] *= 1.1      # bond disorder
    from matjes_amd import llg
    with llg.Hamiltonian(c["lat"].dim_lat, c["lat"].periodic) as h:
        h.add_coo(1, 1, r, cc, v)
        h.add_coo(3, 1, *H.coo_zeeman())
        h.set_bfield(H.b_field_sites())
        h.set_state(c["m0"])
        B = h.get_eff_field()
        h.steps(50, 2, 1.0, 0.1)
        m = h.get_state()

    import types
    Hx = types.SimpleNamespace(lat=H.lat, zeeman=H.zeeman, b_field_sites=H.b_field_sites, dipolar=None,
                               coo_mm=lambda: (r, cc, v))
    o = oracle_lib.Oracle(Hx)
    assert np.max(np.abs(B - o.field(c["m0"]))) <= 1e-14 * np.max(np.abs(B))
    assert np.max(np.abs(m - o.steps(c["m0"], 50, 2, 1.0, 0.1))) <= 1e-12


def test_coo_with_a_moved_entry_is_not_compressed(oracle_lib):
    """A COO whose entry counts per (offset, component) are right but whose
    sites are not (one bond entry removed at one site and duplicated at
    another) is not a uniform stencil: it must take the generic path and give
    the field of exactly those entries."""
    from matjes_amd import llg, workloads
    c = _case(dict(workloads.SQUARE_JDKB, size=[12, 10, 1], seed=5))
    H = c["H"]
    r, cc, v = H.coo_mm()
    r, cc, v = r.copy(), cc.copy(), v.copy()
    # the entry coupling site 3 to its +x neighbour, components (0, 0) (1-based
    # rows 3 i + a + 1): moved to site 17 and its +x neighbour
    n = H.lat.dim_lat[0]
    k_from = [k for k in range(len(r)) if r[k] == 3 * 3 + 1 and cc[k] == 3 * 4 + 1][0]
    r[k_from], cc[k_from] = 3 * 17 + 1, 3 * 18 + 1
    assert 17 % n != n - 1
    with llg.Hamiltonian(c["lat"].dim_lat, c["lat"].periodic) as h:
        h.add_coo(1, 1, r, cc, v)
        h.add_coo(3, 1, *H.coo_zeeman())
        h.set_bfield(H.b_field_sites())
        h.set_state(c["m0"])
        B = h.get_eff_field()
    import types
    Hx = types.SimpleNamespace(lat=H.lat, zeeman=H.zeeman, b_field_sites=H.b_field_sites, dipolar=None,
                               coo_mm=lambda: (r, cc, v))
    ref = oracle_lib.Oracle(Hx).field(c["m0"])
    assert np.max(np.abs(B - ref)) <= 1e-14 * np.max(np.abs(ref))
    # and it differs from the uniform stencil's field at the two sites
    assert np.max(np.abs(ref - oracle_lib.Oracle(H).field(c["m0"]))) > 1e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("integ", [6, 2])
def test_dipolar_film_trajectory_vs_oracle(oracle_lib, integ):
    from matjes_amd import workloads
    s = dict(workloads.FILM_DIP, size=[8, 8, 4], periodic=[True, True, False], seed=12, init="tilted_z")
    c = _case(s)
    o = oracle_lib.Oracle(c["H"])
    mo, Eo = o.steps(c["m0"], 40, integ, 1.0, 0.1, energies=True)
    with _dev(c) as h:
        Eg = h.steps(40, integ, 1.0, 0.1, energies=True)
        mg = h.get_state()
    assert np.max(np.abs(mg - mo)) <= 1e-12
    assert np.max(np.abs(Eg - Eo)) <= 1e-12 * max(np.max(np.abs(Eo)), 1e-3)


@pytest.mark.parametrize("layers", ["1", "0"])
def test_dipolar_padded_trajectory_vs_oracle(oracle_lib, layers):
    """Open boundaries (zero padding in every direction): the input grid's
    padding is zeroed once at set-up and only the lattice's sites are written
    per field evaluation, so every later field must still be the oracle's:
    40 Depondt steps and the field of the state they reach, in the film form
    (2D transforms and layer kernel) and with the 3D transform."""
    from matjes_amd import workloads
    s = dict(workloads.FILM_DIP, size=[10, 8, 3], periodic=[False, False, False], seed=21, init="tilted_z")
    c = _case(s)
    o = oracle_lib.Oracle(c["H"])
    mo, Eo = o.steps(c["m0"], 40, 6, 1.0, 0.1, energies=True)
    old = os.environ.get("MJ_DIP_LAYERS")
    os.environ["MJ_DIP_LAYERS"] = layers
    try:
        with _dev(c) as h:
            Eg = h.steps(40, 6, 1.0, 0.1, energies=True)
            mg = h.get_state()
            B = h.get_eff_field()
    finally:
        if old is None:
            os.environ.pop("MJ_DIP_LAYERS", None)
        else:
            os.environ["MJ_DIP_LAYERS"] = old
    assert np.max(np.abs(mg - mo)) <= 1e-12
    assert np.max(np.abs(Eg - Eo)) <= 1e-12 * max(np.max(np.abs(Eo)), 1e-3)
    Bo = o.field(mg)
    assert np.max(np.abs(B - Bo)) <= 1e-13 * np.max(np.abs(Bo))


@pytest.mark.parametrize("integ", [6, 2])
def test_dipolar_film_compiled_kernels_vs_oracle(oracle_lib, integ):
    """The film workload's kernel path: above 4096 sites the per-Hamiltonian
    stage kernels read the dipolar field of the FFT grid (Depondt, and Heun
    with the thermal field drawn in-kernel) on a 32x32x6 film periodic in
    plane; field, energies and 40 steps against the oracle."""
    from matjes_amd import workloads
    s = dict(workloads.FILM_DIP, size=[32, 32, 6], periodic=[True, True, False], seed=16, init="tilted_z")
    c = _case(s)
    assert c["lat"].ncell >= 4096
    o = oracle_lib.Oracle(c["H"])
    kT = 10.0 * oracle_lib.K_B if integ == 2 else 0.0
    mo, Eo = o.steps(c["m0"], 40, integ, 1.0, 0.1, kT=kT, seed=5, energies=True)
    with _dev(c, "stencil") as h:
        B = h.get_eff_field()
        Eg = h.steps(40, integ, 1.0, 0.1, kT=kT, seed=5, energies=True)
        mg = h.get_state()
    Bo = o.field(c["m0"])
    assert np.max(np.abs(B - Bo)) <= 1e-13 * np.max(np.abs(Bo))
    assert np.max(np.abs(mg - mo)) <= 1e-12
    assert np.max(np.abs(Eg - Eo)) <= 1e-12 * max(np.max(np.abs(Eo)), 1e-3)


def test_full_size_properties():
    """1024x1024 triangular, Heun at T=0: |m| kept at 1, energy does not rise with damping."""
    from matjes_amd import workloads
    c = _case(dict(workloads.TRIANGULAR_J3DKB, size=[1024, 1024, 1], seed=1))
    with _dev(c, "stencil") as h:
        E = h.steps(20, 2, 0.5, 0.05, energies=True)
        m = h.get_state()
    assert np.max(np.abs(np.linalg.norm(m, axis=1) - 1.0)) < 1e-13
    assert np.all(np.diff(E) <= 1e-12 * np.abs(E[:-1]))


def test_full_size_translation_covariance():
    """Shifting the initial state by lattice vectors shifts the result (periodic 1024x1024)."""
    from matjes_amd import workloads
    c = _case(dict(workloads.TRIANGULAR_J3DKB, size=[1024, 1024, 1], seed=1))
    m0 = c["m0"].reshape(1024, 1024, 3)
    sh = np.roll(np.roll(m0, 37, axis=0), 301, axis=1).reshape(-1, 3)
    with _dev(c, "stencil") as h:
        h.steps(10, 2, 0.5, 0.05)
        a = h.get_state().reshape(1024, 1024, 3)
        h.set_state(sh)
        h.steps(10, 2, 0.5, 0.05)
        b = h.get_state().reshape(1024, 1024, 3)
    assert np.array_equal(np.roll(np.roll(a, 37, axis=0), 301, axis=1), b)


def test_error_paths():
    from matjes_amd import _lib, llg
    with llg.Hamiltonian((8, 8, 1), (True, True, False)) as h:
        with pytest.raises(_lib.MatjesError):
            h.set_state(np.zeros((7, 3)))
        with pytest.raises(_lib.MatjesError):
            h.steps(1, 2, 1.0, 0.1)           # no state
        h.set_state(np.tile([0.0, 0.0, 1.0], (64, 1)))
        with pytest.raises(_lib.MatjesError):
            h.steps(1, 2, 1.0, 0.1)           # no Hamiltonian
        with pytest.raises(_lib.MatjesError):
            h.steps(1, 3, 1.0, 0.1)           # unknown integrator
        with pytest.raises(_lib.MatjesError):
            h.add_coo(1, 1, [1, 400], [2, 3], [1.0, 1.0])   # index out of range


@pytest.mark.parametrize("name", ["traj_sq8_heun_update_time", "traj_sq8_euler_update_time"])
def test_update_time_vs_reference(name):
    """Adaptive timestep (update_time.f90:41-75) against the reference's own run."""
    d = golden(name)
    c = _case(d["setup"])
    s = d["setup"]
    n = int(d["steps"][-1])
    with _dev(c, "stencil") as h:
        ts_next, used = h.steps_adaptive(50, s["integration"], s["dt"], s["alpha"], s["discretization"])
        m50 = h.get_state()
        _, used2 = h.steps_adaptive(n - 50, s["integration"], ts_next, s["alpha"], s["discretization"], step0=50)
        m = h.get_state()
    ts = np.concatenate([used, used2])
    rel = np.abs(ts - d["timesteps"]) / d["timesteps"]
    assert np.max(rel[:50]) <= 1e-13 and np.max(rel) <= 1e-10
    assert np.max(np.abs(m50 - d["m"][0])) <= 1e-12
    assert np.max(np.abs(m - d["m"][-1])) <= 1e-9


@pytest.mark.parametrize("integ,T", [(2, 10.0), (5, 0.0), (6, 10.0)])
def test_update_time_vs_oracle(oracle_lib, integ, T):
    from matjes_amd import workloads
    c = _case(dict(workloads.TRIANGULAR_J3DKB, size=[20, 16, 1], seed=5))
    kT = T * oracle_lib.K_B
    ts_o = np.zeros(40)
    mo = oracle_lib.Oracle(c["H"]).steps(c["m0"], 40, integ, 0.5, 0.05, kT=kT, seed=9, discretization=40.0,
                                         timesteps=ts_o)
    with _dev(c, "stencil") as h:
        _, ts = h.steps_adaptive(40, integ, 0.5, 0.05, 40.0, kT=kT, seed=9)
        m = h.get_state()
    assert np.max(np.abs(ts - ts_o) / ts_o) <= 1e-12
    assert np.max(np.abs(m - mo)) <= 1e-11


def test_update_time_zero_field_fails():
    from matjes_amd import _lib, llg
    from matjes_amd import lattice as L
    lat = L.Lattice(np.eye(3), (1.0, 1.0, 1.0), (4, 4, 1), (True, True, False), [("Fe", (0, 0, 0), 1.0)])
    H = L.Hamiltonian(lat)
    H.add_zeeman([0.0, 0.0, 0.0])
    with llg.Hamiltonian(lat.dim_lat, lat.periodic) as h:
        h.load_terms(H)
        h.set_state(np.tile([1.0, 0.0, 0.0], (16, 1)))
        with pytest.raises(_lib.MatjesError):
            h.steps_adaptive(2, llg.EULER, 1.0, 0.1, 15.0)


def test_spindynamics_outputs(oracle_lib, tmp_path):
    """EM.dat rows after steps j with mod(j-1, Efreq) == 0 and magnetic_<tag>.dat
    snapshots (spindynamics.f90:336-409), against the oracle."""
    from matjes_amd import emio, llg, workloads
    c = _case(dict(workloads.SQUARE_JDKB, size=[12, 10, 1], seed=6))
    with _dev(c, "stencil") as h:
        rec = llg.spindynamics(h, 25, llg.HEUN, 1.0, 0.1, efreq=10, outdir=str(tmp_path), gra_freq=12,
                               H_ext=(0.0, 0.0, 0.5))
    assert [r[0] for r in rec] == [1, 11, 21]
    o = oracle_lib.Oracle(c["H"])
    m, last = c["m0"], 0
    for r in rec:
        m, E = o.steps(m, r[0] - last, 2, 1.0, 0.1, energies=True)
        last = r[0]
        assert abs(r[2] - E[-1]) <= 1e-12 * abs(E[-1])
        assert r[1] == (r[0] - 1) * 1.0
    em = np.loadtxt(tmp_path / "EM.dat")
    assert em.shape == (3, 19) and list(em[:, 0]) == [1, 11, 21]
    np.testing.assert_allclose(em[:, 2], [r[2] for r in rec], rtol=1e-11)
    # snapshots at j = 1, 13, 25 -> tags 0, 1, 2
    m25 = o.steps(c["m0"], 25, 2, 1.0, 0.1)
    snap = emio.read_config(emio.config_name(2, str(tmp_path)))
    np.testing.assert_allclose(snap, m25, rtol=1e-13, atol=1e-15)
    assert os.path.exists(emio.config_name(0, str(tmp_path))) and os.path.exists(emio.config_name(1, str(tmp_path)))


@pytest.mark.parametrize("size", [(40, 36), (96, 80)])
@pytest.mark.parametrize("integ,T", [(1, 10.0), (2, 10.0), (5, 10.0), (6, 0.0)])
def test_graph_replay_equals_eager(integ, T, size):
    """Chunks of 32 steps replay a captured graph (device step counter, staged
    energies); states and every step's energy equal the launch-by-launch path,
    on the ahead-of-time kernels (40x36) and on the per-Hamiltonian compiled
    ones with the direct HBM-to-LDS fill (96x80; integration 2 as the fused
    step)."""
    from matjes_amd import workloads
    c = _case(dict(workloads.TRIANGULAR_J3DKB, size=[size[0], size[1], 1], seed=8))
    kT = T * 0.000086173324
    res = []
    for g in ("1", "0"):
        old = os.environ.get("MJ_GRAPH")
        os.environ["MJ_GRAPH"] = g
        try:
            with _dev(c, "stencil") as h:
                E1 = h.steps(100, integ, 0.5, 0.05, kT=kT, seed=4, step0=7, energies=True)
                E2 = h.steps(70, integ, 0.5, 0.05, kT=kT, seed=4, step0=107, energies=True)
                res.append((E1, E2, h.get_state()))
        finally:
            if old is None:
                os.environ.pop("MJ_GRAPH", None)
            else:
                os.environ["MJ_GRAPH"] = old
    for x, y in zip(res[0], res[1]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("size,periodic", [([16, 12, 4], [True, True, False]), ([10, 8, 3], [False, False, False]),
                                           ([12, 12, 1], [True, True, False]), ([8, 8, 6], [True, True, True]),
                                           ([12, 10, 8], [True, True, False])])
def test_dipolar_film_form_equals_3d_transform(oracle_lib, size, periodic):
    """Films (<= 32 layers) use 2D transforms and layer mixing in k-space; the
    field equals the 3D transform's and the oracle's."""
    from matjes_amd import workloads
    c = _case(dict(workloads.FILM_DIP, size=size, periodic=periodic, seed=12))
    out = {}
    for v in ("1", "0"):
        old = os.environ.get("MJ_DIP_LAYERS")
        os.environ["MJ_DIP_LAYERS"] = v
        try:
            with _dev(c, "stencil") as h:
                B = h.get_eff_field()
                E = h.steps(20, 6, 1.0, 0.1, energies=True)
                out[v] = (B, E, h.get_state())
        finally:
            if old is None:
                os.environ.pop("MJ_DIP_LAYERS", None)
            else:
                os.environ["MJ_DIP_LAYERS"] = old
    ref = oracle_lib.Oracle(c["H"]).field(c["m0"])
    for v in ("1", "0"):
        assert np.max(np.abs(out[v][0] - ref)) <= 1e-13 * np.max(np.abs(ref))
    assert np.max(np.abs(out["1"][2] - out["0"][2])) <= 1e-12
    assert np.max(np.abs(out["1"][1] - out["0"][1])) <= 1e-12 * np.max(np.abs(out["0"][1]))


@pytest.mark.parametrize("size,periodic", [([16, 12, 4], [True, True, False]), ([10, 8, 3], [False, False, False]),
                                           ([64, 64, 6], [True, True, False]), ([32, 25, 16], [True, True, False]),
                                           ([20, 13, 32], [False, False, False]), ([16, 12, 2], [True, True, False])])
def test_dipolar_z_parity_form_equals_full_tensor(size, periodic):
    """The layer kernel with half of the dipolar tensor (K(q)(-dz) taken as
    +-K(q)(dz), checked exactly at set-up; the dipolar form at the film's
    compiled layer count) against the general kernel with the full tensor:
    field, energies and a Depondt trajectory bit for bit (64x64x6: the
    per-Hamiltonian compiled stage kernels; 2, 16 and 32 layers: k points
    past the last block's end)."""
    from matjes_amd import workloads
    c = _case(dict(workloads.FILM_DIP, size=size, periodic=periodic, seed=13))
    out = {}
    for v in ("1", "0"):
        old = os.environ.get("MJ_DIP_PARITY")
        os.environ["MJ_DIP_PARITY"] = v
        try:
            with _dev(c, "stencil") as h:
                B = h.get_eff_field()
                E = h.steps(30, 6, 1.0, 0.1, energies=True)
                out[v] = (B, E, h.get_state())
        finally:
            if old is None:
                os.environ.pop("MJ_DIP_PARITY", None)
            else:
                os.environ["MJ_DIP_PARITY"] = old
    for x, y in zip(out["1"], out["0"]):
        assert np.array_equal(x, y)


def test_deterministic_sib_64x64_10k_steps_vs_oracle(oracle_lib):
    """The 64x64 deterministic workload at full length: square lattice, J 1 meV,
    interfacial D 0.3 meV, K 0.05 meV along z, B 0.5 T, SIB dt 1 fs, alpha 0.1,
    T = 0, 10,000 steps, energy recorded every 100 steps; device against the
    oracle's SIB (8 OpenMP threads), 1e-10 per spin and in energy."""
    from matjes_amd import workloads
    c = _case(dict(workloads.SQUARE_JDKB, size=[64, 64, 1], seed=0))
    lib = oracle_lib.lib()
    lib.or_set_num_threads(8)
    try:
        o = oracle_lib.Oracle(c["H"])
        mo, Eo = c["m0"], []
        for _ in range(100):
            mo, E = o.steps(mo, 100, 5, 1.0, 0.1, energies=True)
            Eo.append(E[-1])
    finally:
        lib.or_set_num_threads(1)
    with _dev(c, "stencil") as h:
        Eg = [h.steps(100, 5, 1.0, 0.1, energies=True)[-1] for _ in range(100)]
        m = h.get_state()
    rel = np.max(np.linalg.norm(m - mo, axis=1) / np.linalg.norm(mo, axis=1))
    assert rel <= 1e-10, rel
    assert np.max(np.abs(np.array(Eg) - np.array(Eo))) <= 1e-10 * np.max(np.abs(Eo))


@pytest.mark.parametrize("name", ["field_sq8", "field_tri8", "field_sq8_bmap"])
def test_term_mult_matches_reference_field(oracle_lib, name):
    """mult_r / mult_l of every registered term: against the terms' own CSR, and
    -(sum_t (H_t m + H_t^T m) + Z^T b) is the reference's B_eff (get_eff_field,
    hamiltonian_collection.f90:433-454)."""
    d = golden(name)
    c = _case(d["setup"])
    H = c["H"]
    n3 = 3 * c["lat"].ncell
    rng = np.random.default_rng(4)
    v = rng.normal(size=n3)
    r0 = rng.normal(size=n3)
    m = d["m0"].reshape(-1)
    B = np.zeros(n3)
    with _dev(c, "coo", resolved=True) as h:
        for t, blocks in enumerate(H.mm_terms.values()):
            M, MT = oracle_lib.csr_pair(*H.coo_mm(blocks=blocks), n3)
            hv = h.term_mult(t, v, side=0).reshape(-1)
            htv = h.term_mult(t, v, side=1, res=r0, alpha=-0.5, beta=2.0).reshape(-1)
            assert np.max(np.abs(hv - M @ v)) <= 1e-14 * np.max(np.abs(M @ v))
            ref = -0.5 * (MT @ v) + 2.0 * r0
            assert np.max(np.abs(htv - ref)) <= 1e-14 * np.max(np.abs(ref))
            B -= h.term_mult(t, m, side=0).reshape(-1)
            B -= h.term_mult(t, m, side=1).reshape(-1)
        nz = len(H.mm_terms)
        bsite = H.b_field_sites().reshape(-1)
        zt = h.term_mult(nz, bsite, side=1).reshape(-1)
        Z = oracle_lib.csr_pair(*H.coo_zeeman(), n3)[0]
        assert np.max(np.abs(zt - Z.T @ bsite)) <= 1e-14 * np.max(np.abs(zt))
        assert np.max(np.abs(h.term_mult(nz, m, side=0).reshape(-1) - Z @ m)) <= 1e-14 * np.max(np.abs(Z @ m))
        B -= zt
        with pytest.raises(Exception):
            h.term_mult(nz + 1, v)
    assert np.max(np.abs(B - d["field"])) <= 1e-14 * np.max(np.abs(d["field"]))
