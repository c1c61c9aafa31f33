"""Batched systems and parallel tempering against the single-system path (MI355X).

A batched system runs the same kernel body with the same Philox counters as
the single-system path, so its trajectory must equal mj_steps bit for bit,
whatever its tile shape and whatever else is in the batch.  Replica-exchange
decisions are restated on the host from the device energies with the oracle's
Philox4x32-10 and must match the device's temperature assignment exactly."""
import numpy as np
import pytest

from matjes_amd import lattice as L
from matjes_amd import workloads

pytestmark = pytest.mark.gpu

DT = 1.0


def _systems():
    """Uneven sizes, open and periodic boundaries, different J/D/B, T > 0."""
    specs = [
        dict(workloads.SQUARE_JDKB, size=[33, 20, 1], seed=1, J=[0.7e-3], D=0.2e-3, B=[0.0, 0.0, 1.5]),
        dict(workloads.SQUARE_JDKB, size=[64, 64, 1], seed=2),
        dict(workloads.SQUARE_JDKB, size=[40, 8, 1], periodic=[False, False, False], seed=3, J=[1.8e-3],
             D=0.5e-3, B=[0.2, 0.0, 0.3]),
        dict(workloads.TRIANGULAR_J3DKB, size=[24, 24, 1], seed=4),
    ]
    temps = [5.0, 10.0, 0.5, 20.0]
    return [(workloads.build_case(s), T) for s, T in zip(specs, temps)]


def _plain(case, nsteps, kT, seed, alpha=0.1, integ=2):
    from matjes_amd import llg
    with llg.from_case(case, form="stencil") as h:
        h.steps(nsteps, integ, DT, alpha, kT=kT, seed=seed)
        return h.get_state(), h.energy()


def _batch(cases, tiles=None, alpha=0.1):
    from matjes_amd import batch, llg
    b = batch.Batch()
    hs = []
    for k, (c, T) in enumerate(cases):
        h = llg.from_case(c, form="stencil")
        b.add(h, alpha, kT=T * L.K_B, seed=100 + k)
        hs.append(h)
    if tiles is not None:
        b.set_tiles(tiles)
    return b, hs


@pytest.mark.parametrize("tiles", [None, [(16, 16)] * 4, [(128, 2), (64, 4), (32, 8), (16, 16)]])
def test_batched_trajectories_equal_single_system(tiles):
    cases = _systems()
    b, hs = _batch(cases, tiles)
    try:
        b.steps(120, DT)            # two captured 50-step segments + 20 eager steps
        E = b.energies()
        for k, (c, T) in enumerate(cases):
            m_ref, e_ref = _plain(c, 120, T * L.K_B, 100 + k)
            assert np.array_equal(b.state(k), m_ref), k
            # mj_energy gives the total, mj_batch_energies the energy per site
            assert abs(E[k] * c["lat"].ncell - e_ref) <= 1e-13 * abs(e_ref)
    finally:
        b.close()
        for h in hs:
            h.close()


@pytest.mark.parametrize("integ,mode", [(1, "FUSED"), (5, "FUSED"), (6, "FUSED"), (5, "STAGES"), (6, "STAGES")])
def test_batched_integrators_equal_single_system(integ, mode):
    """Integrations 1 (Euler: out of place, the buffers alternating inside a
    captured segment), 5 (SIB) and 6 (Depondt), with the thermal field in
    their first stage, SIB and Depondt as one fused launch per step (the
    default; the ring draws its thermal field too) or one launch per stage:
    every system equals mj_steps bit for bit after two captured 50-step
    segments and 20 eager steps."""
    from matjes_amd import batch
    cases = _systems()
    b, hs = _batch(cases)
    try:
        b.set_launch(getattr(batch, mode))
        b.steps(120, DT, integrator=integ)
        E = b.energies()
        for k, (c, T) in enumerate(cases):
            m_ref, e_ref = _plain(c, 120, T * L.K_B, 100 + k, integ=integ)
            assert np.array_equal(b.state(k), m_ref), (integ, k)
            assert abs(E[k] * c["lat"].ncell - e_ref) <= 1e-13 * abs(e_ref)
    finally:
        b.close()
        for h in hs:
            h.close()


@pytest.mark.parametrize("integ", [1, 2, 5, 6])
def test_batched_site_b_maps_equal_single_system(integ):
    """Systems with per-site B maps (the Zeeman block applied to each site's
    own B) beside a uniform-field system: every system equals mj_steps bit for
    bit after two captured segments and 20 eager steps; replicas that share
    one map run parallel tempering with it."""
    from matjes_amd import batch, llg
    specs = [
        dict(workloads.SQUARE_JDKB, size=[33, 20, 1], seed=1, bmap_sigma=0.3),
        dict(workloads.SQUARE_JDKB, size=[48, 40, 1], seed=2),
        dict(workloads.TRIANGULAR_J3DKB, size=[24, 24, 1], seed=4, bmap_sigma=0.2),
    ]
    cases = [(workloads.build_case(s), T) for s, T in zip(specs, [5.0, 10.0, 20.0])]
    assert cases[0][0]["bmap"] is not None and cases[1][0]["bmap"] is None
    b, hs = _batch(cases)
    try:
        b.steps(120, DT, integrator=integ)
        E = b.energies()
        for k, (c, T) in enumerate(cases):
            m_ref, e_ref = _plain(c, 120, T * L.K_B, 100 + k, integ=integ)
            assert np.array_equal(b.state(k), m_ref), (integ, k)
            assert abs(E[k] * c["lat"].ncell - e_ref) <= 1e-13 * abs(e_ref)
    finally:
        b.close()
        for h in hs:
            h.close()
    c = cases[0][0]
    ladder = np.array([1.0, 3.0]) * L.K_B
    b = batch.Batch()
    hs = []
    try:
        for r in range(2):
            h = llg.from_case(c, form="stencil")
            b.add(h, 0.1, kT=ladder[r], seed=40 + r)
            hs.append(h)
        b.set_pt(0, 2, ladder, 1000, seed=1)
        b.steps(60, DT, integrator=integ)
        for r in range(2):
            m_ref, _ = _plain(c, 60, ladder[r], 40 + r, integ=integ)
            assert np.array_equal(b.state(r), m_ref)
    finally:
        b.close()
        for h in hs:
            h.close()


@pytest.mark.parametrize("integ", [1, 2, 5, 6])
def test_batched_dipolar_systems_equal_single_system(integ):
    """Films with the FFT dipolar term (in-plane periodic and open, the layer
    transforms; a thicker open film through the 3D transform) beside a
    stencil-only system: each dipolar system's H_fft is formed by the batch's
    own copy of its transform before every stage, and every system equals
    mj_steps bit for bit (every step eager, one launch per stage)."""
    import os
    specs = [
        dict(workloads.FILM_DIP, size=[16, 12, 4], periodic=[True, True, False], seed=12),
        dict(workloads.SQUARE_JDKB, size=[33, 20, 1], seed=1),
        dict(workloads.FILM_DIP, size=[10, 8, 3], periodic=[False, False, False], seed=13),
    ]
    cases = [(workloads.build_case(s), T) for s, T in zip(specs, [5.0, 10.0, 20.0])]
    old = os.environ.get("MJ_DIP_LAYERS")
    try:
        for layers in ("1", "0"):
            os.environ["MJ_DIP_LAYERS"] = layers
            b, hs = _batch(cases)
            try:
                b.steps(30, DT, integrator=integ)
                E = b.energies()
                for k, (c, T) in enumerate(cases):
                    m_ref, e_ref = _plain(c, 30, T * L.K_B, 100 + k, integ=integ)
                    assert np.array_equal(b.state(k), m_ref), (layers, k)
                    assert abs(E[k] * c["lat"].ncell - e_ref) <= 1e-13 * abs(e_ref)
            finally:
                b.close()
                for h in hs:
                    h.close()
    finally:
        if old is None:
            os.environ.pop("MJ_DIP_LAYERS", None)
        else:
            os.environ["MJ_DIP_LAYERS"] = old


@pytest.mark.parametrize("mode", ["STAGES", "PERSISTENT", "PERSISTENT_FUSED"])
@pytest.mark.parametrize("tiles", [None, [(128, 2), (64, 4), (32, 8), (16, 16)]])
def test_launch_forms_equal_single_system(mode, tiles):
    """Integration 2's captured segments with one launch per stage, and as one
    persistent launch each (grid barriers between the stages, or between
    fused steps): every system equals mj_steps bit for bit after two segments
    and 20 eager steps.  (The fused per-step form is the default, which the
    other tests here run.)"""
    from matjes_amd import batch
    cases = _systems()
    b, hs = _batch(cases, tiles)
    try:
        b.set_launch(getattr(batch, mode))
        b.steps(120, DT)
        E = b.energies()
        for k, (c, T) in enumerate(cases):
            m_ref, e_ref = _plain(c, 120, T * L.K_B, 100 + k)
            assert np.array_equal(b.state(k), m_ref), k
            assert abs(E[k] * c["lat"].ncell - e_ref) <= 1e-13 * abs(e_ref)
    finally:
        b.close()
        for h in hs:
            h.close()


@pytest.mark.parametrize("mode", ["PERSISTENT", "PERSISTENT_FUSED"])
def test_persistent_barrier_gives_up_and_recovers(mode):
    """A barrier allowed no wait (spin 0) gives up at once: every workgroup
    leaves, the run raises, and the reset barrier serves the next run."""
    from matjes_amd import _lib, batch
    cases = _systems()[:2]
    b, hs = _batch(cases)
    try:
        b.set_launch(getattr(batch, mode), spin=0)
        with pytest.raises(_lib.MatjesError, match="gave up"):
            b.steps(50, DT)
        b.set_launch(getattr(batch, mode))
        b.steps(50, DT)
        assert np.all(np.isfinite(b.energies()))
    finally:
        b.close()
        for h in hs:
            h.close()


def test_unaligned_calls_continue_the_same_trajectory():
    cases = _systems()[:2]
    b, hs = _batch(cases)
    try:
        for n in (7, 50, 13, 50):
            b.steps(n, DT)
        for k, (c, T) in enumerate(cases):
            m_ref, _ = _plain(c, 120, T * L.K_B, 100 + k)
            assert np.array_equal(b.state(k), m_ref)
    finally:
        b.close()
        for h in hs:
            h.close()


def _philox_u(orc, t, e, seed):
    w = orc.philox([t, e & 0xFFFFFFFF, e >> 32, 2], [seed & 0xFFFFFFFF, seed >> 32])
    return float(((w[1] << 32) | w[0]) >> 11) * 2.0 ** -53


@pytest.mark.parametrize("integ,every,mode", [(2, 10, "FUSED"), (2, 10, "STAGES"), (2, 10, "PERSISTENT"),
                                              (2, 10, "PERSISTENT_FUSED"), (1, 5, "STAGES"), (6, 10, "STAGES"),
                                              (6, 10, "FUSED"), (5, 10, "FUSED")])
def test_replica_exchange_matches_host_restatement(oracle_lib, integ, every, mode):
    """Exchange decisions of Heun, Euler and Depondt replicas against the host
    restatement with the oracle's Philox.  The entering energies come from the
    stage-1 kernel (Heun, per stage), or from a field pass before the step
    (fused Heun; Euler and Depondt, whose first stage is noisy); persistent
    segments form them and exchange inside the launch."""
    from matjes_amd import batch, llg
    R, seed = 6, 77
    s = dict(workloads.SQUARE_JDKB, size=[16, 16, 1], seed=9)
    c = workloads.build_case(s)
    temps = np.geomspace(1.0, 100.0, R)
    ladder = temps * L.K_B
    b = batch.Batch()
    hs = []
    try:
        for r in range(R):
            h = llg.from_case(c, form="stencil")
            b.add(h, 0.1, kT=ladder[r], seed=1000 + r)
            hs.append(h)
        b.set_pt(0, R, ladder, every, seed=seed)
        b.set_launch(getattr(batch, mode))
        rung = np.arange(R)            # rung of each replica
        rep = np.arange(R)             # replica on each rung
        acc = np.zeros(R, dtype=np.int64)
        b.steps(every, DT, integrator=integ)
        beta = 1.0 / ladder
        n = 256
        for e in range(1, 9):
            Etot = b.energies() * n
            for t in range(e & 1, R - 1, 2):
                ra, rb = rep[t], rep[t + 1]
                d = (beta[t] - beta[t + 1]) * (Etot[ra] - Etot[rb])
                if d >= 0.0 or _philox_u(oracle_lib, t, e, seed) < np.exp(d):
                    rep[t], rep[t + 1] = rb, ra
                    rung[ra], rung[rb] = t + 1, t
                    acc[t] += 1
            b.steps(every, DT, integrator=integ)
            got, gacc, gatt = b.pt_state(R)
            assert np.array_equal(got, rung), (e, got, rung)
            assert np.array_equal(gacc[:R - 1].astype(np.int64), acc[:R - 1])
        assert acc.sum() > 0
        # states stay finite and on the unit sphere
        for r in range(R):
            m = b.state(r)
            assert np.max(np.abs(np.linalg.norm(m, axis=1) - 1.0)) < 1e-12
    finally:
        b.close()
        for h in hs:
            h.close()


@pytest.mark.parametrize("integ,mode", [(1, "STAGES"), (2, "FUSED"), (2, "PERSISTENT_FUSED"), (5, "FUSED")])
def test_odd_interval_segments_with_two_exchanges_equal_eager_steps(integ, mode):
    """Euler and the fused forms of Heun with an odd exchange interval (5): a
    captured segment holds two intervals (the state buffers alternate an even
    number of times) and its second exchange reads its own step; the same
    replicas stepped one call per step (eager, one launch per stage) end with
    the same states and temperature assignment."""
    from matjes_amd import batch, llg
    s = dict(workloads.SQUARE_JDKB, size=[16, 12, 1], seed=4)
    c = workloads.build_case(s)
    ladder = np.geomspace(1.0, 50.0, 4) * L.K_B
    out = []
    for chunk in (20, 1):
        b = batch.Batch()
        hs = []
        try:
            for r in range(4):
                h = llg.from_case(c, form="stencil")
                b.add(h, 0.1, kT=ladder[r], seed=60 + r)
                hs.append(h)
            b.set_pt(0, 4, ladder, 5, seed=3)
            b.set_launch(getattr(batch, mode))
            for _ in range(40 // chunk):
                b.steps(chunk, DT, integrator=integ)
            tor, acc, att = b.pt_state(4)
            out.append(([b.state(r) for r in range(4)], tor, acc, att))
        finally:
            b.close()
            for h in hs:
                h.close()
    (ms, tor, acc, att), (ms2, tor2, acc2, att2) = out
    assert np.array_equal(tor, tor2) and np.array_equal(acc, acc2) and np.array_equal(att, att2)
    assert att.sum() > 0
    for r in range(4):
        assert np.array_equal(ms[r], ms2[r]), r


def test_replicas_without_swaps_equal_single_system():
    """Exchange every 1000 steps within a 100-step run: no swap happens, so each
    replica is the single-system trajectory at its own temperature."""
    from matjes_amd import batch, llg
    s = dict(workloads.SQUARE_JDKB, size=[20, 12, 1], seed=2)
    c = workloads.build_case(s)
    ladder = np.array([1.0, 3.0, 9.0]) * L.K_B
    b = batch.Batch()
    hs = []
    try:
        for r in range(3):
            h = llg.from_case(c, form="stencil")
            b.add(h, 0.1, kT=ladder[r], seed=40 + r)
            hs.append(h)
        b.set_pt(0, 3, ladder, 1000, seed=1)
        b.steps(100, DT)
        for r in range(3):
            m_ref, _ = _plain(c, 100, ladder[r], 40 + r)
            assert np.array_equal(b.state(r), m_ref)
    finally:
        b.close()
        for h in hs:
            h.close()


def test_autotuner_picks_valid_shapes_and_caches(tmp_path, monkeypatch):
    from matjes_amd import autotune, batch
    monkeypatch.setenv("MJ_TUNE_CACHE", str(tmp_path / "tiles.json"))
    cases = _systems()
    b, hs = _batch(cases)
    try:
        dims = [c["lat"].dim_lat for c, _ in cases]
        choice, rep = autotune.tune(b, dims, DT, reach=2, seg=10)
        assert not rep["cached"] and rep["ms_per_step"] > 0
        assert rep["launch"] in autotune.FORMS and set(rep["ms_per_step_by_form"]) == {"fused", "stages"}
        assert rep["ms_per_step"] == min(rep["ms_per_step_by_form"].values())
        assert b.launch == autotune.FORMS[rep["launch"]]
        for shp in choice.values():
            assert tuple(shp) in autotune.SHAPES
        _, _, tiles = b.info()
        assert [tuple(t) for t in tiles] == autotune.tiles_for(dims, choice)
        b.set_launch(batch.STAGES)
        choice2, rep2 = autotune.tune(b, dims, DT, reach=2, seg=10)
        assert rep2["cached"] and {k: list(v) for k, v in choice.items()} == choice2
        assert rep2["launch"] == rep["launch"] and b.launch == autotune.FORMS[rep["launch"]]
    finally:
        b.close()
        for h in hs:
            h.close()


def test_batch_errors():
    from matjes_amd import _lib, batch, llg
    c = workloads.build_case(dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=1))
    b = batch.Batch()
    try:
        with llg.from_case(c, form="stencil") as h:
            b.add(h, 0.1, kT=0.0)
        with pytest.raises(_lib.MatjesError):
            b.set_tiles([(10, 10)])
        with pytest.raises(_lib.MatjesError):      # integration 4 does not run batched
            b.steps(1, DT, integrator=llg.RK)
        with pytest.raises(_lib.MatjesError):
            b.set_pt(0, 2, [1e-4, 2e-4], 10)
        cd = workloads.build_case(dict(workloads.FILM_DIP, size=[4, 4, 2], seed=1))
        for _ in range(2):                          # dipolar systems run batched, but not as replicas
            with llg.from_case(cd, form="stencil") as hd:
                b.add(hd, 0.1)
        with pytest.raises(_lib.MatjesError):
            b.set_pt(1, 2, [1e-4, 2e-4], 10)
        for sd in (1, 2):                           # replicas need one B map: these two differ
            cm = workloads.build_case(dict(workloads.SQUARE_JDKB, size=[8, 8, 1], seed=sd, bmap_sigma=0.3))
            with llg.from_case(cm, form="stencil") as hm:
                b.add(hm, 0.1)
        with pytest.raises(_lib.MatjesError):
            b.set_pt(3, 2, [1e-4, 2e-4], 10)
    finally:
        b.close()


def test_full_composition_of_the_batched_workload():
    """The batched workload at its real composition: 64 parallel-tempering
    replicas of a 256x256 lattice on a geometric 1..100 K ladder exchanging
    every 50 steps, plus 192 independent lattices with sides log-uniform in
    [32, 512] and their own J, D/J, B at 10 K (12.4M sites), 100 steps.  Every
    spin keeps |m| = 1, energies are finite, swaps are accepted at a rate
    strictly between 0 and 1, and independent systems of the smallest, a
    middle and the largest size equal the single-system run bit for bit."""
    from matjes_amd import batch, llg
    rng = np.random.default_rng(11)
    specs = [(dict(workloads.SQUARE_JDKB, size=[256, 256, 1], seed=1000 + r), None) for r in range(64)]
    for k in range(192):
        side = int(round(np.exp(rng.uniform(np.log(32), np.log(512)))))
        J = rng.uniform(0.5e-3, 2e-3)
        specs.append((dict(workloads.SQUARE_JDKB, size=[side, side, 1], J=[J], D=rng.uniform(0.0, 0.5) * J,
                           B=[0.0, 0.0, rng.uniform(0.0, 2.0)], seed=2000 + k), 10.0 * L.K_B))
    ladder = np.geomspace(1.0, 100.0, 64) * L.K_B
    b = batch.Batch()
    cases = []
    try:
        for k, (s, kT) in enumerate(specs):
            c = workloads.build_case(s)
            cases.append(c)
            with llg.from_case(c, form="stencil") as h:
                b.add(h, 0.1, kT=ladder[k] if kT is None else kT, seed=5000 + k)
        b.set_pt(0, 64, ladder, 50, seed=11)
        sites, _, _ = b.info()
        assert sites > 12_000_000
        b.steps(100, DT)
        E = b.energies()
        assert np.all(np.isfinite(E))
        rung, acc, att = b.pt_state(64)
        assert sorted(rung.tolist()) == list(range(64))
        rate = float(acc[:63].sum()) / float(att[:63].sum())
        assert 0.0 < rate < 1.0
        sides = [specs[k][0]["size"][0] for k in range(64, 256)]
        order = np.argsort(sides)
        sample = [64 + int(order[0]), 64 + int(order[len(order) // 2]), 64 + int(order[-1])]
        for k in sample + [0, 63]:
            m = b.state(k)
            assert np.max(np.abs(np.linalg.norm(m, axis=1) - 1.0)) < 1e-13
        for k in sample:
            m_ref, _ = _plain(cases[k], 100, 10.0 * L.K_B, 5000 + k)
            assert np.array_equal(b.state(k), m_ref), (k, specs[k][0]["size"])
    finally:
        b.close()


def test_full_composition_launch_forms_agree():
    """The batched workload (as above, 100 steps: two segments with an exchange
    each) in all four launch forms: every system's state and the temperature
    assignment are equal bit for bit."""
    from matjes_amd import batch, llg
    rng = np.random.default_rng(11)
    specs = [(dict(workloads.SQUARE_JDKB, size=[256, 256, 1], seed=1000 + r), None) for r in range(64)]
    for k in range(192):
        side = int(round(np.exp(rng.uniform(np.log(32), np.log(512)))))
        J = rng.uniform(0.5e-3, 2e-3)
        specs.append((dict(workloads.SQUARE_JDKB, size=[side, side, 1], J=[J], D=rng.uniform(0.0, 0.5) * J,
                           B=[0.0, 0.0, rng.uniform(0.0, 2.0)], seed=2000 + k), 10.0 * L.K_B))
    ladder = np.geomspace(1.0, 100.0, 64) * L.K_B
    modes = [batch.STAGES, batch.FUSED, batch.PERSISTENT, batch.PERSISTENT_FUSED]
    bs = [batch.Batch() for _ in modes]
    try:
        for k, (s, kT) in enumerate(specs):
            c = workloads.build_case(s)
            for b in bs:
                with llg.from_case(c, form="stencil") as h:
                    b.add(h, 0.1, kT=ladder[k] if kT is None else kT, seed=5000 + k)
        for b, mode in zip(bs, modes):
            b.set_pt(0, 64, ladder, 50, seed=11)
            b.set_launch(mode)
            b.steps(100, DT)
        r0, a0, t0 = bs[0].pt_state(64)
        assert a0[:63].sum() > 0
        for b in bs[1:]:
            r1, a1, t1 = b.pt_state(64)
            assert np.array_equal(r0, r1) and np.array_equal(a0, a1) and np.array_equal(t0, t1)
        for k in range(len(specs)):
            m0 = bs[0].state(k)
            for i in range(1, len(bs)):
                assert np.array_equal(m0, bs[i].state(k)), (k, modes[i])
    finally:
        for b in bs:
            b.close()
