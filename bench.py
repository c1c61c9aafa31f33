#!/usr/bin/env python3
"""Throughput of the FP64 LLG step on one MI355X.

Default workload: the 1024x1024 triangular lattice (exchange J1..J3, interfacial
DMI, easy-axis anisotropy, Zeeman), the reference's two-half-step Heun
(`integration 2`) with dt = 0.5 fs, alpha = 0.05 and a T = 10 K Langevin field
drawn in-kernel from Philox4x32-10, with the total energy evaluated after
every step as the reference's loop does.  A step is one pass of that loop over
the resident state; inputs are in HBM before the timed region starts.

Prints one JSON line with `roofline` (the dominant kernel: the fused step
kernel, which runs both stages of integration 2; its average duration from HIP
events around a block of back-to-back launches on the library's stream, with
energies on as in the timed steps) and `cpu_baseline` (the oracle's C/OpenMP
restatement of the reference loop, timed on a bounded sample of the same
workload on this host).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "site-steps/s, FP64 LLG 1024² (+% HBM peak); batched/adjoint vs plain"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per site per launch (FP64 SoA vectors streamed once), T > 0:
#   stage 1 (HEUN1): read m0, write m1, D1                       = 3 x 24 B
#   stage 2 (HEUN2): read m1, D1, m0 (Langevin term), write m2    = 4 x 24 B
# the fused step kernel does both stages: 168 B of the two-stage algorithm per
# site-step (it moves m0 in and m2 out, 48 B, and keeps m1 and D1 on chip)
BYTES_STAGE = {0: 72.0, 1: 96.0}
BYTES_STEP = 168.0


def workload_setup(size, seed=1):
    from matjes_amd import workloads
    return dict(workloads.TRIANGULAR_J3DKB, size=[size, size, 1], seed=seed)


def cpu_baseline(sample_size, budget_s, threads=None):
    """Oracle (C restatement of the reference loop, OpenMP) on a sample of the workload."""
    from matjes_amd import workloads
    from oracle import oracle as orc
    if threads is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orc.lib().or_set_num_threads(threads)
    c = workloads.build_case(workload_setup(sample_size))
    o = orc.Oracle(c["H"])
    kT = 10.0 * orc.K_B
    m = c["m0"]
    m = o.steps(m, 1, 2, 0.5, 0.05, kT=kT, seed=7, energies=False)     # warm-up
    steps, t0 = 0, time.perf_counter()
    while True:
        m, E = o.steps(m, 1, 2, 0.5, 0.05, kT=kT, seed=7, step0=steps + 1, energies=True)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    orc.lib().or_set_num_threads(1)
    n = c["lat"].ncell
    return {"value": n * steps / el, "unit": "site-steps/s", "cores": threads, "kind": "port",
            "sample": "%dx%d triangular lattice, same Hamiltonian/integrator/thermal field, %d steps "
                      "in %.1f s (CSR H.m + H^T.m per stage, energy every step)" % (sample_size, sample_size,
                                                                                      steps, el)}


def cpu_baseline_mc(sample_size, budget_s, threads=None):
    """Oracle Metropolis sweeps (C, OpenMP over the sites of a colour class) on a
    sample of the Monte Carlo form's workload."""
    from matjes_amd import workloads
    from oracle import oracle as orc
    if threads is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orc.lib().or_set_num_threads(threads)
    c = workloads.build_case(workload_setup(sample_size))
    o = orc.Oracle(c["H"])
    kT = 10.0 * orc.K_B
    st = np.array([0.5, np.cos(0.5), 0.0, 0.0, 0.0])
    m, _, _ = o.mc_sweeps(c["m0"], 1, 0, kT, seed=1, state=st)
    sweeps, t0 = 0, time.perf_counter()
    while True:
        m, _, _ = o.mc_sweeps(m, 1, 0, kT, seed=1, sweep0=sweeps + 1, state=st)
        sweeps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    orc.lib().or_set_num_threads(1)
    n = c["lat"].ncell
    return {"value": n * sweeps / el, "unit": "site-updates/s", "cores": threads, "kind": "port",
            "sample": "%dx%d triangular lattice, same Hamiltonian, sphere proposals at 10 K, %d sweeps in %.1f s "
                      "(CSR rows per single-site energy)" % (sample_size, sample_size, sweeps, el)}


def integrators_form(steps=200, size=1024, seed=1):
    """The plain step's lattice, Hamiltonian and 10 K thermal field with each
    integrator: reference Euler (integration 1), reference two-half-step Heun
    (2), SIB (5) and Depondt (6), dt 0.5 fs, energy every step."""
    import torch
    from matjes_amd import llg, workloads
    from matjes_amd.lattice import K_B
    c = workloads.build_case(workload_setup(size))
    n = c["lat"].ncell
    out = {}
    for name, integ in (("euler", 1), ("heun", 2), ("sib", 5), ("depondt", 6)):
        with llg.from_case(c, form="stencil") as h:
            h.steps(20, integ, 0.5, 0.05, kT=10.0 * K_B, seed=seed, energies=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            h.steps(steps, integ, 0.5, 0.05, kT=10.0 * K_B, seed=seed, step0=20, energies=True)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
        out[name] = {"integration": integ, "site_steps_per_s": n * steps / el, "ms_per_step": el / steps * 1e3}
    return out


def mc_form(sweeps, size=1024, seed=1):
    """Metropolis sweeps of the plain step's lattice and Hamiltonian (triangular
    1024x1024, J1-3 + DMI + K + B) at 10 K: sphereft proposals in the adaptive
    cone, and Ising flips; one pass per colour class (8 classes here)."""
    from matjes_amd import _lib, llg, workloads
    from matjes_amd.lattice import K_B
    c = workloads.build_case(workload_setup(size))
    n = c["lat"].ncell
    res = {}
    with llg.from_case(c, form="stencil") as h:
        for name, smp in (("sphere", _lib.MJ_MC_SPHERE), ("ising", _lib.MJ_MC_ISING)):
            h.set_state(c["m0"])
            h.mc_set_cone(0.5)
            h.mc_sweeps(10, smp, 10.0 * K_B, seed=seed)
            t0 = time.perf_counter()
            h.mc_sweeps(sweeps, smp, 10.0 * K_B, seed=seed, sweep0=10)
            el = time.perf_counter() - t0
            st = h.mc_stats()
            res[name] = {"site_updates_per_s": n * sweeps / el, "ms_per_sweep": el / sweeps * 1e3, "sweeps": sweeps,
                         "colours": st["colouring"][0], "acceptance": st["rate"], "cone": st["cone"],
                         # state read and written once per sweep (48 B per site update)
                         "algorithmic_GBps": 48.0 * n * sweeps / el / 1e9}
        # parallel tempering: 8 replicas on a geometric 2..60 K ladder, exchanges every 5 sweeps
        nrep = 8
        h.set_state(c["m0"])
        h.mc_pt_setup(np.geomspace(2.0, 60.0, nrep) * K_B, cone=0.5)
        h.mc_pt_sweeps(10, _lib.MJ_MC_SPHERE, seed=seed, exchange_every=5)
        t0 = time.perf_counter()
        h.mc_pt_sweeps(sweeps, _lib.MJ_MC_SPHERE, seed=seed, sweep0=10, exchange_every=5)
        el = time.perf_counter() - t0
        ps = h.mc_pt_state()
        res["parallel_tempering"] = {"replicas": nrep, "site_updates_per_s": nrep * n * sweeps / el,
                                     "ms_per_sweep_all_replicas": el / sweeps * 1e3, "sweeps": sweeps,
                                     "swap_acceptance": (ps["swaps_accepted"] / np.maximum(ps["swaps_attempted"], 1)).tolist(),
                                     "algorithmic_GBps": 48.0 * nrep * n * sweeps / el / 1e9}
    return res


def batched_form(steps, seed=11, log=None):
    """64 parallel-tempering replicas of a 256x256 square lattice (J 1 meV, DMI 0.3,
    K 0.05 easy z, B 0.5 T) on a geometric 1..100 K ladder exchanging every 50 steps,
    plus 192 independent square lattices (side log-uniform in [32, 512], J in
    [0.5, 2] meV, D/J in [0, 0.5], B in [0, 2] T along z, T = 10 K); reference Heun
    dt 1 fs, alpha 0.1; graph-captured 50-step segments, the launch form (one
    fused launch per step for the whole batch, or one launch per stage) and the
    tile shapes per size class from the autotuner.  The other launch forms
    (including one persistent launch per segment with per-stage or fused
    tiles) are timed on the same batch after it."""
    from matjes_amd import autotune, batch, llg, workloads
    from matjes_amd.lattice import K_B
    rng = np.random.default_rng(seed)
    specs = []
    for r in range(64):
        specs.append((dict(workloads.SQUARE_JDKB, size=[256, 256, 1], seed=1000 + r), None))
    for k in range(192):
        L = int(round(np.exp(rng.uniform(np.log(32), np.log(512)))))
        J = rng.uniform(0.5e-3, 2e-3)
        D = rng.uniform(0.0, 0.5) * J
        B = rng.uniform(0.0, 2.0)
        specs.append((dict(workloads.SQUARE_JDKB, size=[L, L, 1], J=[J], D=D, B=[0.0, 0.0, B], seed=2000 + k),
                      10.0 * K_B))
    ladder = np.geomspace(1.0, 100.0, 64) * K_B
    dims = [tuple(s["size"]) for s, _ in specs]

    def build():
        b = batch.Batch()
        for k, (s, kT) in enumerate(specs):
            c = workloads.build_case(s)
            with llg.from_case(c, form="stencil") as h:
                b.add(h, 0.1, kT=ladder[k] if kT is None else kT, seed=5000 + k)
        b.set_pt(0, 64, ladder, 50, seed=seed)
        return b

    t0 = time.perf_counter()
    b = build()
    t_build = time.perf_counter() - t0
    choice, rep = autotune.tune(b, dims, 1.0, reach=1, log=log)
    b.close()
    b = build()                      # fresh states; tuned form and shapes
    form = getattr(batch, rep["launch"].upper())
    b.set_launch(form)
    b.set_tiles(autotune.tiles_for(dims, choice))
    sites, wg, _ = b.info()
    b.steps(50, 1.0)                 # warm-up: graph capture, first replay
    ms = b.profile(steps, 1.0)
    rung, acc, att = b.pt_state(64)
    E = b.energies()
    forms = {}
    for name in ("FUSED", "STAGES", "PERSISTENT", "PERSISTENT_FUSED"):
        if getattr(batch, name) == form:
            continue
        b.set_launch(getattr(batch, name))
        b.steps(50, 1.0)
        forms[name.lower()] = b.profile(steps, 1.0) / steps
    b.close()
    acc_rate = float(acc[:63].sum()) / max(float(att[:63].sum()), 1.0)
    return {"site_steps_per_s": sites * steps / (ms * 1e-3), "ms_per_step": ms / steps, "sites": int(sites),
            "systems": len(specs), "steps": steps, "workgroups_per_stage": int(wg),
            "tiles_by_size_class": {k: list(v) for k, v in choice.items()}, "autotune": rep,
            "pt_swap_acceptance": acc_rate, "energies_finite": bool(np.all(np.isfinite(E))),
            "build_s": t_build, "launch": rep["launch"], "ms_per_step_other_launch_forms": forms}


def adjoint_form(steps=2000, ckpt=100, seed=3):
    """128x128 square lattice, J/D/K/B as above plus a per-site field map
    (Gaussian, sigma 0.05 meV / (mu_s mu_B) in Tesla); SIB dt 1 fs alpha 0.2, N
    steps checkpointed every 100; loss sum |m_N - m_skyrmion|^2 (Neel skyrmion of
    radius 10 sites, theta = pi (1 - r/R)).  Gradients w.r.t. (J, D, K, B_z) are compared with central
    differences (h = 1e-6 relative) of the plain single-system SIB on the GPU."""
    import torch
    from matjes_amd import adjoint, llg, workloads
    from matjes_amd import lattice as Lt
    lat = workloads.make_lattice(dict(workloads.SQUARE_JDKB, size=[128, 128, 1]))
    n = lat.ncell
    rng = np.random.default_rng(seed)
    hmap = (0.05e-3 / (2.0 * Lt.MU_B)) * rng.normal(size=(n, 3))
    m0 = workloads.random_sphere(n, seed)
    cx = cy = 63.5
    yy, xx = np.meshgrid(np.arange(128.0), np.arange(128.0), indexing="ij")
    r = np.hypot(xx - cx, yy - cy).ravel()
    th = np.where(r < 10.0, np.pi * (1.0 - r / 10.0), 0.0)      # theta = pi (1 - r/R), R = 10 sites
    ph = np.arctan2((yy - cy).ravel(), (xx - cx).ravel())
    tgt = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    theta = np.array([1e-3, 0.3e-3, -0.05e-3])
    bext = np.array([0.0, 0.0, 0.5])
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    hd, m0d, tg = dev(hmap), dev(m0), dev(tgt)
    run = adjoint.SIBRunner(lat, ("J1", "D", "K"), nsteps=steps, timestep=1.0, damping=0.2, ckpt_every=ckpt)
    out = torch.empty_like(m0d)
    run.forward(theta, bext, hd, m0d, out)            # warm-up
    run.backward((2 * (out - tg)).contiguous(), torch.empty_like(m0d), None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run.forward(theta, bext, hd, m0d, out)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ghm = torch.empty_like(m0d)
    gth, gb = run.backward((2 * (out - tg)).contiguous(), ghm, None)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    mN = out.cpu().numpy()
    run.close()

    def plain_loss(th_, b_):
        H = Lt.Hamiltonian(lat)
        H.add_exchange(1, th_[0])
        H.add_dmi(1, th_[1], "Ir")
        H.add_anisotropy([0.0, 0.0, 1.0], th_[2])
        H.add_zeeman(np.asarray(b_)[None, :] + hmap)
        old = os.environ.get("MJ_JIT")
        os.environ["MJ_JIT"] = "0"     # ahead-of-time kernels (bit-identical; no compile per perturbed theta)
        try:
            with llg.Hamiltonian(lat.dim_lat, lat.periodic) as d:
                d.load_terms(H, form="stencil")
                d.set_state(m0)
                d.steps(steps, llg.SIB, 1.0, 0.2)
                m = d.get_state()
        finally:
            if old is None:
                os.environ.pop("MJ_JIT", None)
            else:
                os.environ["MJ_JIT"] = old
        return float(np.sum((m - tgt) ** 2)), m
    L0, mp = plain_loss(theta, bext)
    # the plain SIB step on the same lattice (the per-Hamiltonian kernels, as
    # the plain form runs): a 16k-site lattice is latency bound in either form
    H = Lt.Hamiltonian(lat)
    H.add_exchange(1, theta[0])
    H.add_dmi(1, theta[1], "Ir")
    H.add_anisotropy([0.0, 0.0, 1.0], theta[2])
    H.add_zeeman(bext[None, :] + hmap)
    with llg.Hamiltonian(lat.dim_lat, lat.periodic) as d:
        d.load_terms(H, form="stencil")
        d.set_state(m0)
        d.steps(200, llg.SIB, 1.0, 0.2)
        tp0 = time.perf_counter()
        d.steps(steps, llg.SIB, 1.0, 0.2)
        tp = time.perf_counter() - tp0
    plain_same = n * steps / tp
    fd, err = [], []
    for k in range(4):
        if k < 3:
            e = 1e-6 * abs(theta[k])
            d = np.eye(3)[k] * e
            g = (plain_loss(theta + d, bext)[0] - plain_loss(theta - d, bext)[0]) / (2 * e)
            ga = gth[k]
        else:
            e = 1e-6 * abs(bext[2])
            d = np.array([0.0, 0.0, e])
            g = (plain_loss(theta, bext + d)[0] - plain_loss(theta, bext - d)[0]) / (2 * e)
            ga = gb[2]
        fd.append(g)
        err.append(abs(ga - g) / abs(g))
    return {"forward_site_steps_per_s": n * steps / (t1 - t0),
            "forward_backward_site_steps_per_s": n * steps / (t2 - t0),
            "forward_ms": (t1 - t0) * 1e3, "backward_ms": (t2 - t1) * 1e3, "sites": n, "steps": steps,
            "checkpoint_every": ckpt, "loss": L0,
            "plain_sib_same_lattice_site_steps_per_s": plain_same,
            "forward_backward_vs_plain_same_lattice": n * steps / (t2 - t0) / plain_same,
            "grad_J_D_K_Bz": [float(gth[0]), float(gth[1]), float(gth[2]), float(gb[2])],
            "central_diff_plain_gpu": fd, "rel_err_vs_central_diff": err,
            "max_abs_state_diff_vs_plain_sib": float(np.max(np.abs(mN - mp)))}


def dipolar_form(steps, warmup=5):
    """Simple-cubic film 256x256x16 (a = 0.4 nm, mu = 2 mu_B, J 2 meV, K 0.2 meV easy z)
    with the full dipole-dipole field by FFT convolution; Depondt dt 1 fs, alpha 0.1,
    from +z with random tilts up to 10 degrees.  Two grids: the reference's
    convention (periodic in plane with supercell images: N_rep 256x256x32) and
    the literal zero-padded 512x512x32 (open in every direction)."""
    from matjes_amd import llg, workloads
    res = {}
    for name, per in (("periodic_inplane_256x256x32", [True, True, False]), ("padded_512x512x32", [False, False, False])):
        t0 = time.perf_counter()
        c = workloads.build_case(dict(workloads.FILM_DIP, size=[256, 256, 16], periodic=per, init="tilted_z",
                                      tilt_deg=10.0, seed=2))
        t_build = time.perf_counter() - t0
        n = c["lat"].ncell
        with llg.from_case(c, form="stencil") as h:
            # energies on in the warm-up as in the timed steps: their kernels
            # are compiled on first use (~0.2 s)
            h.steps(warmup, llg.DEPONDT, 1.0, 0.1, energies=True)
            tot, kms, kcnt = h.profile_steps(steps, llg.DEPONDT, 1.0, 0.1, energies=True)
            E = h.energy()
        res[name] = {"site_steps_per_s": n * steps / (tot * 1e-3), "ms_per_step": tot / steps, "sites": n,
                     "N_rep": list(c["H"].dipolar["N_rep"]), "steps": steps,
                     "stage_kernels_ms_per_step": (kms[0] + kms[1]) / steps,
                     "fft_and_launch_gaps_ms_per_step": (tot - kms[0] - kms[1]) / steps, "energy_per_site": E, "operator_build_s": t_build}
        print("dipolar %s done" % name, file=sys.stderr, flush=True)
    return res


def host_cpu():
    """Model and logical CPU count of the host the CPU baseline ran on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(),
            "available_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def load_traffic(key, size):
    """HBM bytes per launch of a kernel ("fused", "stage1", "stage2") from the
    committed rocprofv3 --pmc summary for this workload size, or None."""
    p = os.path.join(REPO, "profiles", "traffic_heun_%d.json" % size)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("%s_hbm_bytes_per_launch" % key)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--device-warmup-s", type=float, default=1.5,
                    help="seconds of untimed steps before the warmup steps (GPU clock ramp)")
    ap.add_argument("--forms", default="batched,adjoint,dipolar,mc,integrators",
                    help="comma list of the forms measured beside the plain step: batched, adjoint, "
                         "dipolar, mc, integrators ('' for none)")
    ap.add_argument("--batched-steps", type=int, default=1000)
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--dipolar-steps", type=int, default=100)
    ap.add_argument("--mc-sweeps", type=int, default=200)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="only run a plain device copy of this many MiB 20 times (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.calibrate_copy:
        import torch
        n = a.calibrate_copy * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(20):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 20}), flush=True)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the stencil would shard "
              "as y-slabs with point-to-point halo exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from matjes_amd import llg, workloads
    import torch
    assert torch.cuda.is_available(), "no GPU visible"

    s = workload_setup(a.size)
    c = workloads.build_case(s)
    n = c["lat"].ncell
    kT = 10.0 * 0.000086173324
    h = llg.from_case(c, form="stencil")
    # the GPU's clocks rise over the first second of sustained load (after 0.5 s
    # a step still takes ~17% longer: 48.9 against 42.2 us after 1, 2 or 3 s)
    # and fall again within milliseconds of idling (an
    # upload of the state between the phases costs up to 15% on the next few
    # hundred steps): the device warm-up, the warmup steps, the timed steps and
    # the per-launch timing run back to back, each on the state the previous
    # one left (the cost of a step does not depend on the state).  The warm-up
    # steps take the timed call's parameters, energies included, so that the
    # graph of 32-step chunks the library replays is captured before the timing
    t_dev = time.perf_counter()
    j = 0
    while time.perf_counter() - t_dev < a.device_warmup_s:
        h.steps(64, llg.HEUN, 0.5, 0.05, kT=kT, seed=1, step0=j, energies=True)
        torch.cuda.synchronize()
        j += 64
    h.steps(a.warmup, llg.HEUN, 0.5, 0.05, kT=kT, seed=1, step0=j, energies=True)
    torch.cuda.synchronize()
    j += a.warmup
    t0 = time.perf_counter()
    E = h.steps(a.steps, llg.HEUN, 0.5, 0.05, kT=kT, seed=1, step0=j, energies=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    j += a.steps
    value = n * a.steps / el
    assert np.all(np.isfinite(E))

    # kernel durations from HIP events on the library's stream, energies on
    tot_ms, kms, kcnt = h.profile_steps(a.profile_steps, llg.HEUN, 0.5, 0.05, kT=kT, seed=2, energies=True)
    avg = {k: kms[k] / max(kcnt[k], 1) for k in (0, 1)}
    step_bytes = (BYTES_STAGE[0] + BYTES_STAGE[1]) * n
    if kcnt[0] == 0:
        # one kernel per step (both stages fused)
        kname, kbytes, tkey = "fused step (integration 2: stage 1 on tile + halo, stage 2, thermal field, energy)", \
            BYTES_STEP * n, "fused"
        dom = 1
    else:
        dom = 1 if kms[1] >= kms[0] else 0          # kernel with the larger share of the step
        kname = "stage 2 (HEUN2 + thermal field)" if dom == 1 else "stage 1 (HEUN1)"
        kbytes, tkey = BYTES_STAGE[dom] * n, "stage%d" % (dom + 1)
    ach = kbytes / (avg[dom] * 1e-3) / 1e9
    traffic = load_traffic(tkey, a.size)   # from the rocprof --pmc summary
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "site-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random initial spins on S^2, seed 1; Philox thermal field)",
        "device_warmup_s": a.device_warmup_s,
        "config": {"workload": "triangular %dx%d, J1-3 + interfacial DMI + K + B, reference Heun "
                               "(integration 2) dt=0.5 fs alpha=0.05 T=10 K, energy every step" % (a.size, a.size),
                   "sites": n, "integrator": "heun(integration 2)", "timestep_fs": 0.5, "damping": 0.05,
                   "T_K": 10.0},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": kbytes,
                     "avg_launch_ms": avg[dom], "launches_timed": int(kcnt[dom]),
                     "stage1_avg_launch_ms": avg[0] if kcnt[0] else None,
                     "step_algorithmic_GBps": step_bytes * value / n / 1e9},
    }
    h.close()
    forms = [f for f in a.forms.split(",") if f]
    if forms:
        out["forms"] = {}
    if "batched" in forms:
        bf = batched_form(a.batched_steps, log=lambda m: print(m, file=sys.stderr, flush=True))
        bf["vs_plain"] = bf["site_steps_per_s"] / value
        bf["vs_plain_note"] = ("plain: 1024x1024 triangular J1-3 lattice, energy every step; batched: square "
                               "J1 lattices, energies only on exchange steps (every 50)")
        out["forms"]["batched"] = bf
        print("batched form done", file=sys.stderr, flush=True)
    if "adjoint" in forms:
        af = adjoint_form(a.adjoint_steps)
        af["forward_backward_vs_plain"] = af["forward_backward_site_steps_per_s"] / value
        out["forms"]["adjoint"] = af
        print("adjoint form done", file=sys.stderr, flush=True)
    if "dipolar" in forms:
        out["forms"]["dipolar"] = dipolar_form(a.dipolar_steps)
    if "integrators" in forms:
        out["forms"]["integrators"] = integrators_form()
        print("integrators form done", file=sys.stderr, flush=True)
    if "mc" in forms:
        out["forms"]["mc"] = mc_form(a.mc_sweeps)
        print("mc form done", file=sys.stderr, flush=True)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.cpu_sample, a.cpu_seconds)
        # the reference's Fortran loop itself is serial: the same sample on one core
        one = cpu_baseline(a.cpu_sample, a.cpu_seconds / 2, threads=1)
        out["cpu_baseline"]["single_core_value"] = one["value"]
        out["cpu_baseline"]["single_core_sample"] = one["sample"]
        out["cpu_baseline"]["host"] = host_cpu()
        out["cpu_baseline"]["plan_deviation"] = ("a 512x512 sample timed for a fixed budget (15 s on all host "
                                                 "threads, 7.5 s on one) instead of the median of 5 x 100 steps "
                                                 "at 1024x1024, to keep the run within minutes; the rate per "
                                                 "site-step of the restated loop does not depend on the size here")
        if "mc" in forms:
            cm = cpu_baseline_mc(256, a.cpu_seconds / 3)
            out["cpu_baseline"]["mc"] = cm
            out["forms"]["mc"]["sphere"]["vs_cpu_port"] = out["forms"]["mc"]["sphere"]["site_updates_per_s"] / cm["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
