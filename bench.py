#!/usr/bin/env python3
"""Throughput of the FP64 LLG step on one MI355X.

Default workload: the 1024x1024 triangular lattice (exchange J1..J3, interfacial
DMI, easy-axis anisotropy, Zeeman), the reference's two-half-step Heun
(`integration 2`) with dt = 0.5 fs, alpha = 0.05 and a T = 10 K Langevin field
drawn in-kernel from Philox4x32-10, with the total energy evaluated after
every step as the reference's loop does.  A step is one pass of that loop over
the resident state; inputs are in HBM before the timed region starts.

Prints one JSON line with `roofline` (dominant kernel: the stage-2 kernel,
per-launch duration from HIP events on the library's stream) and
`cpu_baseline` (the oracle's C/OpenMP restatement of the reference loop, timed
on a bounded sample of the same workload on this host).
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
# algorithmic bytes per site per launch (FP64 SoA state streamed once):
#   stage 1 (HEUN1): read m0, write m1, D1                     = 3 x 24 B
#   stage 2 (HEUN2): read m1, D1, m0 (thermal field), write m2  = 4 x 24 B
BYTES_STAGE = {0: 72.0, 1: 96.0}


def workload_setup(size, seed=1):
    from matjes_amd import workloads
    return dict(workloads.TRIANGULAR_J3DKB, size=[size, size, 1], seed=seed)


def cpu_baseline(sample_size, budget_s):
    """Oracle (C restatement of the reference loop, OpenMP) on a sample of the workload."""
    from matjes_amd import workloads
    from oracle import oracle as orc
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


def load_traffic(kernel, size):
    p = os.path.join(REPO, "profiles", "traffic_%s_%d.json" % (kernel, size))
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
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
    ap.add_argument("--profile-steps", type=int, default=20)
    a = ap.parse_args()
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
    h.steps(a.warmup, llg.HEUN, 0.5, 0.05, kT=kT, seed=1, step0=0, energies=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E = h.steps(a.steps, llg.HEUN, 0.5, 0.05, kT=kT, seed=1, step0=a.warmup, energies=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = n * a.steps / el
    assert np.all(np.isfinite(E))

    # per-kernel launch durations from HIP events on the library's stream
    tot_ms, kms, kcnt = h.profile_steps(a.profile_steps, llg.HEUN, 0.5, 0.05, kT=kT, seed=2)
    avg = {k: kms[k] / max(kcnt[k], 1) for k in (0, 1)}
    dom = 1 if kms[1] >= kms[0] else 0          # kernel with the larger share of the step
    ach = BYTES_STAGE[dom] * n / (avg[dom] * 1e-3) / 1e9
    step_bytes = (BYTES_STAGE[0] + BYTES_STAGE[1]) * n
    traffic = load_traffic("heun2" if dom == 1 else "heun1", a.size)
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
        "config": {"workload": "triangular %dx%d, J1-3 + interfacial DMI + K + B, reference Heun "
                               "(integration 2) dt=0.5 fs alpha=0.05 T=10 K, energy every step" % (a.size, a.size),
                   "sites": n, "integrator": "heun(integration 2)", "timestep_fs": 0.5, "damping": 0.05,
                   "T_K": 10.0},
        "roofline": {"bound": "hbm", "kernel": "k_stage<HEUN2,noise>" if dom == 1 else "k_stage<HEUN1>", "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": BYTES_STAGE[dom] * n,
                     "avg_launch_ms": avg[dom], "stage1_avg_launch_ms": avg[0],
                     "step_algorithmic_GBps": step_bytes * value / n / 1e9},
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.cpu_sample, a.cpu_seconds)
    h.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
