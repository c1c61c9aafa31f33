"""The t_H_hip backend (matjes_amd/fortran/ham_hip.f90) inside the reference's
own Fortran class tree: oracle/_ref/mjref_hipham builds the reference's
resolved Hamiltonian with its dense backend, rebuilds every term as a t_H_hip
through the reference's init_coo, moves them into a reference `hamiltonian`
collection and compares the collection's get_eff_field / energy /
energy_resolved and each term's mult_r / mult_l on both backends; the
library's step loop run from the reference's own setup against the
reference's trajectories; and the reference's own factory (get_Htype /
get_Htype_N with `Hamiltonian_mode 8` in its input) allocating t_H_hip for
every term, under the reference's field, minimiser and GNEB."""
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, REPO, golden

pytestmark = pytest.mark.gpu

EXE = os.path.join(REPO, "oracle", "_ref", "mjref_hipham")


@pytest.mark.skipif(not os.path.isfile(EXE), reason="reference build with the hip backend not present")
@pytest.mark.parametrize("name", ["field_sq8", "field_tri8", "field_sq8_bmap"])
def test_hip_backend_in_the_reference_collection(name, tmp_path):
    sys.path.insert(0, GOLDEN)
    import gen_golden
    from matjes_amd import workloads
    d = golden(name)
    setup = d["setup"]
    case = workloads.build_case(setup)
    (tmp_path / "input").write_text(gen_golden.input_text(setup, "field"))
    np.savetxt(tmp_path / "magnetic.inp", case["m0"], fmt="%.17e")
    if case.get("bmap") is not None:
        np.savetxt(tmp_path / "Bfield.inp", case["bmap"], fmt="%.17e")
    r = subprocess.run(["bash", "-c", "ulimit -s unlimited && %s" % EXE], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    v = np.loadtxt(tmp_path / "hipham.out")
    n3 = 3 * case["lat"].ncell
    nH = (len(v) - 2 - 2 * n3) // (2 + 4 * n3)
    assert len(v) == 2 + 2 * nH + 2 * n3 + 4 * nH * n3 and nH == len(d["term_names"])
    e_dense, e_hip = v[0], v[1]
    assert abs(e_hip - e_dense) <= 1e-13 * abs(e_dense)
    assert abs(e_dense - d["energy"]) <= 1e-13 * abs(d["energy"])
    terms = v[2:2 + 2 * nH].reshape(nH, 2)
    assert np.max(np.abs(terms[:, 1] - terms[:, 0])) <= 1e-13 * np.max(np.abs(terms[:, 0]))
    B = v[2 + 2 * nH:2 + 2 * nH + 2 * n3].reshape(2, n3)
    assert np.max(np.abs(B[1] - B[0])) <= 1e-14 * np.max(np.abs(B[0]))
    assert np.max(np.abs(B[1] - d["field"])) <= 1e-14 * np.max(np.abs(d["field"]))
    prods = v[2 + 2 * nH + 2 * n3:].reshape(nH, 4, n3)
    for t in range(nH):
        for dense, hip in ((prods[t, 0], prods[t, 1]), (prods[t, 2], prods[t, 3])):
            assert np.max(np.abs(hip - dense)) <= 1e-14 * np.max(np.abs(dense)), (name, t)
    # copy + add_child of the M-M terms (combine_Hamiltonians) against the dense terms' sum
    comb = np.loadtxt(tmp_path / "hipham_comb.out").reshape(2, n3)
    assert np.max(np.abs(comb[0] - comb[1])) <= 1e-14 * np.max(np.abs(comb[1]))


@pytest.mark.skipif(not os.path.isfile(EXE), reason="reference build with the hip backend not present")
@pytest.mark.parametrize("name", ["traj_sq8_euler", "traj_sq8_heun", "traj_tri8_heun_T10", "traj_tri8_euler_T10",
                                  "traj_sq8_heun_bmap", "traj_sq8_rk4", "traj_tri8_rk3"])
def test_library_step_loop_in_the_reference_setup(name, tmp_path):
    """`mjref_hipham traj`: the reference's setup_simu builds the lattice and the
    combined Hamiltonian from its own input file, and the library runs the steps
    (mj_steps over the reference's COO entries, B order parameter and state;
    at finite T the in-kernel thermal field) in place of the reference's stage
    loop; the output steps against the reference's own trajectory."""
    sys.path.insert(0, GOLDEN)
    import gen_golden
    from matjes_amd import workloads
    d = golden(name)
    s = d["setup"]
    case = workloads.build_case(s)
    (tmp_path / "input").write_text(gen_golden.input_text(s, "traj"))
    np.savetxt(tmp_path / "magnetic.inp", case["m0"], fmt="%.17e")
    if case.get("bmap") is not None:
        np.savetxt(tmp_path / "Bfield.inp", case["bmap"], fmt="%.17e")
    if s.get("T", 0.0) > 0.0:
        # the library draws the thermal field in-kernel; the reference scaled the
        # same normals by the sigma it computed itself
        (tmp_path / "noise_seed").write_text("%d\n" % s["noise_seed"])
    r = subprocess.run(["bash", "-c", "ulimit -s unlimited && %s traj %d" % (EXE, s["integration"])],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    n3 = 3 * case["lat"].ncell
    rec = np.loadtxt(tmp_path / "traj.out").reshape(-1, n3 + 2)
    assert np.array_equal(rec[:, 0].astype(np.int64), d["steps"])
    for k in range(len(rec)):
        assert np.max(np.abs(rec[k, 2:].reshape(-1, 3) - d["m"][k])) <= 1e-12, k
        assert abs(rec[k, 1] - d["E"][k]) <= 1e-12 * abs(d["E"][k]), k


@pytest.mark.skipif(not os.path.isfile(EXE), reason="reference build with the hip backend not present")
@pytest.mark.parametrize("name", ["field_sq8", "field_tri8", "field_dip_open", "field_dip_ttf"])
def test_fft_backend_with_the_reference_operator_builders(name, tmp_path, oracle_lib):
    """`mjref_hipham fft`: the reference's own FFT operator builders (J, DMI,
    anisotropy, dipolar) fill fft_H_hip terms (matjes_amd/fortran/
    fft_ham_hip.f90) and the library transforms them.  Each term's energy and
    field against the reference's dense term of the same description (the
    dipolar only with every direction open: periodic directions sum other
    images in the two forms, dipolar_magnetic.f90:58-177 against 181-273), and
    the collection holding the FFT terms beside the other terms against the
    oracle's field and energy (whose dipolar follows get_dipolar_fft)."""
    sys.path.insert(0, GOLDEN)
    import gen_golden
    from matjes_amd import workloads
    d = golden(name)
    setup = d["setup"]
    case = workloads.build_case(setup)
    (tmp_path / "input").write_text(gen_golden.input_text(setup, "field"))
    np.savetxt(tmp_path / "magnetic.inp", case["m0"], fmt="%.17e")
    r = subprocess.run(["bash", "-c", "ulimit -s unlimited && %s fft" % EXE], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    v = np.loadtxt(tmp_path / "fft.out")
    descs = (tmp_path / "fft_desc.out").read_text().splitlines()
    n3 = 3 * case["lat"].ncell
    nf, pos = int(v[0]), 1
    assert nf == len(descs) and nf >= 2
    open_all = not any(setup["periodic"])
    for t in range(nf):
        jd, Ef, Ed = int(v[pos]), v[pos + 1], v[pos + 2]
        Bf, Bd = v[pos + 3:pos + 3 + n3], v[pos + 3 + n3:pos + 3 + 2 * n3]
        pos += 3 + 2 * n3
        assert jd > 0, descs[t]
        if descs[t] == "magnetic dipolar" and not open_all:
            continue
        scale = np.max(np.abs(Bd))
        assert np.max(np.abs(Bf - Bd)) <= 1e-12 * scale, (descs[t], np.max(np.abs(Bf - Bd)) / scale)
        assert abs(Ef - Ed) <= 1e-12 * (abs(Ed) + 0.5 * scale * case["lat"].ncell), descs[t]
    Emix, Eres = v[pos], v[pos + 1]
    Bmix, Bres = v[pos + 2:pos + 2 + n3], v[pos + 2 + n3:pos + 2 + 2 * n3]
    assert pos + 2 + 2 * n3 == len(v)
    o = oracle_lib.Oracle(case["H"])
    ref = o.field(case["m0"]).reshape(-1)
    assert np.max(np.abs(Bmix - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert abs(Emix - o.energy(case["m0"])) <= 1e-12 * abs(Emix)
    if open_all or not setup.get("dipolar"):
        assert np.max(np.abs(Bmix - Bres)) <= 1e-12 * np.max(np.abs(Bres))
        assert abs(Emix - Eres) <= 1e-12 * abs(Eres)
    # copy + add (combine_Hamiltonians): one operator, the sum of the fields
    comb = np.loadtxt(tmp_path / "fft_comb.out").reshape(2, n3)
    assert np.max(np.abs(comb[0] - comb[1])) <= 1e-12 * np.max(np.abs(comb[1]))


@pytest.mark.skipif(not os.path.isfile(EXE), reason="reference build with the hip backend not present")
@pytest.mark.parametrize("name", ["field_sq8", "field_tri8"])
def test_reference_minimizer_on_the_hip_collection(name, tmp_path):
    """`mjref_hipham minimize`: the reference's own minimiser (minimize_run, which
    evaluates get_eff_field every iteration) from the same state on the dense
    collection and on the t_H_hip collection: the energy falls and both end in
    the same state."""
    sys.path.insert(0, GOLDEN)
    import gen_golden
    from matjes_amd import workloads
    d = golden(name)
    setup = d["setup"]
    case = workloads.build_case(setup)
    text = gen_golden.input_text(setup, "field")
    text += "steps 300\nmasse 1.0\ntimestep 0.1\nmin_Efreq 1000000\nconvergence_criteria 1e-14\n"
    (tmp_path / "input").write_text(text)
    np.savetxt(tmp_path / "magnetic.inp", case["m0"], fmt="%.17e")
    r = subprocess.run(["bash", "-c", "ulimit -s unlimited && %s minimize" % EXE], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    v = np.loadtxt(tmp_path / "minimize.out")
    n3 = 3 * case["lat"].ncell
    assert len(v) == 2 + 2 * n3
    E_dense, E_hip = v[0], v[1]
    assert E_dense < d["energy"] - 0.1 * abs(d["energy"]), (E_dense, d["energy"])
    assert abs(E_hip - E_dense) <= 1e-10 * abs(E_dense)
    assert np.max(np.abs(v[2 + n3:] - v[2:2 + n3])) <= 1e-10


def _gneb_inputs(d, n=20, itrmax=500, ham_mode=5):
    """An n-spin open chain (J 1 meV, easy axis z) reversing between two tilted
    states in 8 images: the reference's own GNEB input keys (GNEB_io.f90)."""
    sys.path.insert(0, GOLDEN)
    import gen_golden
    s = dict(geometry="square", periodic=[False, False, False], moment=1.0, J=[1e-3], D=0.0, K=-3e-4,
             K_axis=[0.0, 0.0, 1.0], B=[0.0, 0.0, 0.0], size=[n, 1, 1])
    text = _with_mode(gen_golden.input_text(s, "field"), ham_mode) + "\n".join([
        "initpath 1", "do_gneb_simple T", "do_gneb_ci F", "nim 8", "mep_itrmax %d" % itrmax, "mep_ftol 1e-9",
        "meptraj_step 100000", "en_zero I", "amp_rnd_path 0.0", "spring 0.5", "vpo_dt 0.01", "vpo_mass 1.0",
        "sample_num 20"]) + "\n"
    i = np.arange(n)
    th = 0.05 + 0.04 * np.sin(0.7 * i)
    mi = np.stack([np.sin(th) * np.cos(0.3 * i), np.sin(th) * np.sin(0.3 * i), np.cos(th)], 1)
    (d / "input").write_text(text)
    np.savetxt(d / "momfile_i", mi, fmt="%.17e")
    np.savetxt(d / "momfile_f", mi * np.array([1.0, 1.0, -1.0]), fmt="%.17e")
    np.savetxt(d / "magnetic.inp", mi, fmt="%.17e")


@pytest.mark.skipif(not os.path.isfile(EXE), reason="reference build with the hip backend not present")
def test_reference_gneb_on_the_hip_collection(tmp_path):
    """`mjref_hipham gneb`: the reference's GNEB (500 VPO iterations of 8 images,
    forces from get_eff_field) on the dense collection and on the t_H_hip
    collection: the same energy path (the reference's en_path.out) and the same
    images."""
    out = {}
    for which in ("dense", "hip"):
        d = tmp_path / which
        d.mkdir()
        _gneb_inputs(d)
        r = subprocess.run(["bash", "-c", "ulimit -s unlimited && %s gneb %s" % (EXE, which)], cwd=d,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        out[which] = (np.loadtxt(d / "en_path.out"), np.loadtxt(d / "image-GNEB_4.dat"))
    (ed, imd), (eh, imh) = out["dense"], out["hip"]
    assert ed.shape == (8, 3) and np.max(ed[:, 1]) > 1e-3           # a barrier, energies relative to image 1
    assert np.max(np.abs(eh - ed)) <= 1e-9 * np.max(np.abs(ed))
    assert np.max(np.abs(imh - imd)) <= 1.5e-8          # image files carry 8 decimals


def _with_mode(text, ham_mode):
    """The input text with the reference's `Hamiltonian_mode` key set: 5 its
    dense backend, 8 t_H_hip (the case the patched factory adds)."""
    out, n = re.subn(r"(?m)^Hamiltonian_mode\s+\d+", "Hamiltonian_mode %d" % ham_mode, text)
    assert n == 1
    return out


def _factory(d, task, *args):
    r = subprocess.run(["bash", "-c", "ulimit -s unlimited && %s factory %s %s" % (EXE, task, " ".join(args))],
                       cwd=d, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    kinds = (d / "factory.out").read_text().split("\n", 1)[0].split()
    values = np.loadtxt(d / "factory.out", skiprows=1) if task != "gneb" else None
    return kinds, values, float(np.loadtxt(d / "timing.out"))


def _field_case(d, setup, ham_mode, extra=""):
    sys.path.insert(0, GOLDEN)
    import gen_golden
    from matjes_amd import workloads
    case = workloads.build_case(setup)
    (d / "input").write_text(_with_mode(gen_golden.input_text(setup, "field"), ham_mode) + extra)
    np.savetxt(d / "magnetic.inp", case["m0"], fmt="%.17e")
    if case.get("bmap") is not None:
        np.savetxt(d / "Bfield.inp", case["bmap"], fmt="%.17e")
    return case


@pytest.mark.skipif(not os.path.isfile(EXE), reason="reference build with the hip backend not present")
@pytest.mark.parametrize("name", ["field_sq8", "field_tri8", "field_sq8_bmap"])
def test_reference_factory_allocates_the_hip_backend(name, tmp_path):
    """`Hamiltonian_mode 8`: the reference's setup_simu, through its own
    get_Htype_N / get_Htype, builds every resolved and combined term as a
    t_H_hip; the combined collection's energy and field and the resolved
    terms' energies equal the reference's golden values."""
    d = golden(name)
    case = _field_case(tmp_path, d["setup"], 8)
    kinds, v, _ = _factory(tmp_path, "field", "2")
    res, comb = " ".join(kinds).split(" | ")
    assert set(res.split()) == {"hip"} and set(comb.split()) == {"hip"}, kinds
    assert len(res.split()) == len(d["term_names"])
    n3 = 3 * case["lat"].ncell
    nH = len(d["term_names"])
    assert len(v) == 1 + nH + n3
    assert abs(v[0] - d["energy"]) <= 1e-13 * abs(d["energy"])
    assert np.max(np.abs(v[1:1 + nH] - d["term_energies"])) <= 1e-13 * np.max(np.abs(d["term_energies"]))
    assert np.max(np.abs(v[1 + nH:] - d["field"])) <= 1e-14 * np.max(np.abs(d["field"]))


MIN_KEYS = "steps %d\nmasse 1.0\ntimestep 0.1\nmin_Efreq 1000000\nconvergence_criteria 1e-30\n"


def _factory_pair(tmp_path, task, setup=None, steps=300, gneb_n=20, gneb_itr=500, nrep="20"):
    """The same task on the reference's dense backend (mode 5) and on t_H_hip
    (mode 8); returns {mode: (kinds, values, seconds)}."""
    out = {}
    for mode in (5, 8):
        d = tmp_path / ("mode%d" % mode)
        d.mkdir(parents=True)
        if task == "gneb":
            _gneb_inputs(d, gneb_n, gneb_itr, mode)
            out[mode] = _factory(d, "gneb") + (np.loadtxt(d / "en_path.out"), np.loadtxt(d / "image-GNEB_4.dat"))
        else:
            _field_case(d, setup, mode, MIN_KEYS % steps if task == "minimize" else "")
            out[mode] = _factory(d, task, nrep)
        want = "hip" if mode == 8 else "dense"
        assert set(out[mode][0]) == {want, "|"}, out[mode][0]
    return out


@pytest.mark.skipif(not os.path.isfile(EXE), reason="reference build with the hip backend not present")
def test_reference_factory_minimizer_and_gneb_equal_the_dense_backend(tmp_path):
    """The reference's minimize_run and GNEB given H_comb as main.f90 gives it:
    built by the factory as t_H_hip (mode 8) they end where the dense backend
    (mode 5) ends."""
    d = golden("field_tri8")
    out = _factory_pair(tmp_path / "min", "minimize", d["setup"])
    (_, vd, _), (_, vh, _) = out[5], out[8]
    assert vd[0] < d["energy"] - 0.1 * abs(d["energy"])
    assert abs(vh[0] - vd[0]) <= 1e-10 * abs(vd[0])
    assert np.max(np.abs(vh[1:] - vd[1:])) <= 1e-10
    out = _factory_pair(tmp_path / "gneb", "gneb")
    (ed, imd), (eh, imh) = out[5][3:], out[8][3:]
    assert ed.shape == (8, 3) and np.max(ed[:, 1]) > 1e-3
    assert np.max(np.abs(eh - ed)) <= 1e-9 * np.max(np.abs(ed))
    assert np.max(np.abs(imh - imd)) <= 1.5e-8


@pytest.mark.skipif(not os.path.isfile(EXE), reason="reference build with the hip backend not present")
def test_reference_factory_iteration_time(tmp_path):
    """Seconds per iteration of the reference's field, minimiser and GNEB on a
    32x32 triangular film (J1-J3, DMI, anisotropy, Zeeman; 3072-row operators)
    and a 1024-spin chain of 8 images (the GNEB time includes its path set-up
    and output, spread over the iterations), dense backend (mode 5) against t_H_hip
    (mode 8), with the results equal; with MJ_FACTORY_TIMING set to a path the
    figures are written there as JSON."""
    d = golden("field_tri8")
    setup = dict(d["setup"], size=[32, 32, 1])
    steps, itr = 20, 10
    fld = _factory_pair(tmp_path / "fld", "field", setup, nrep="5")
    mini = _factory_pair(tmp_path / "min", "minimize", setup, steps=steps)
    gneb = _factory_pair(tmp_path / "gneb", "gneb", gneb_n=1024, gneb_itr=itr)
    assert np.max(np.abs(fld[8][1] - fld[5][1])) <= 1e-13 * np.max(np.abs(fld[5][1]))
    assert np.max(np.abs(mini[8][1] - mini[5][1])) <= 1e-10 * np.max(np.abs(mini[5][1]))
    assert np.max(np.abs(gneb[8][3] - gneb[5][3])) <= 1e-9 * np.max(np.abs(gneb[5][3]))
    rep = {"lattice": "triangular 32x32, 1024 sites; GNEB chain 1024 spins x 8 images",
           "field_s_per_call": {"dense": fld[5][2], "hip": fld[8][2]},
           "minimizer_s_per_iteration": {"dense": mini[5][2] / steps, "hip": mini[8][2] / steps},
           "gneb_s_per_iteration": {"dense": gneb[5][2] / itr, "hip": gneb[8][2] / itr}}
    print(json.dumps(rep))
    if os.environ.get("MJ_FACTORY_TIMING"):
        with open(os.environ["MJ_FACTORY_TIMING"], "w") as f:
            json.dump(rep, f, indent=1)
    assert rep["minimizer_s_per_iteration"]["hip"] < rep["minimizer_s_per_iteration"]["dense"]
