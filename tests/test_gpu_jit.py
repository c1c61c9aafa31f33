"""Per-Hamiltonian compiled stage kernels against the ahead-of-time kernels (MI355X).

Both run the same kernel body with the same summation order, so states,
fields and energies must agree bit for bit."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(case, jit, integ, steps, kT=0.0, tile=None):
    from matjes_amd import llg
    old = {k: os.environ.get(k) for k in ("MJ_JIT",)}
    os.environ["MJ_JIT"] = "1" if jit else "0"
    try:
        with llg.from_case(case, form="stencil", tile=tile) as h:
            B = h.get_eff_field()
            E = h.steps(steps, integ, 0.5, 0.05, kT=kT, seed=3, energies=True)
            m = h.get_state()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return B, E, m


@pytest.mark.parametrize("integ,T", [(1, 0.0), (2, 0.0), (2, 10.0), (5, 10.0), (6, 0.0)])
@pytest.mark.parametrize("which", ["tri", "sq_bmap", "film"])
def test_jit_equals_aot(integ, T, which):
    from matjes_amd import workloads
    if which == "tri":
        s = dict(workloads.TRIANGULAR_J3DKB, size=[96, 80, 1], seed=1)
    elif which == "sq_bmap":
        s = dict(workloads.SQUARE_JDKB, size=[64, 72, 1], seed=2, bmap_sigma=0.3)
    else:
        s = dict(workloads.FILM_DIP, size=[32, 32, 6], seed=3, dipolar=False)
    c = workloads.build_case(s)
    kT = T * 0.000086173324
    a = _run(c, False, integ, 30, kT)
    b = _run(c, True, integ, 30, kT)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("tile", [(64, 4, 1), (32, 8, 1), (128, 2, 1), (64, 8, 1), (16, 4, 4)])
def test_jit_tiles(tile):
    from matjes_amd import workloads
    s = dict(workloads.TRIANGULAR_J3DKB, size=[256, 192, 1], seed=4)
    if tile[2] > 1:
        s = dict(workloads.FILM_DIP, size=[64, 48, 8], seed=4, dipolar=False)
    c = workloads.build_case(s)
    a = _run(c, False, 2, 10, 10.0 * 0.000086173324, tile=tile)
    b = _run(c, True, 2, 10, 10.0 * 0.000086173324, tile=tile)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def _run_fuse(case, fuse, steps, kT=0.0, tile=None, chunks=(1,), spt=1, integ=2):
    from matjes_amd import llg
    # MJ_GRAPH=1: chunks of 64 and more replay the captured graph on both paths
    env = {"MJ_FUSE": "1" if fuse else "0", "MJ_FUSE_SPT": str(spt), "MJ_GRAPH": "1"}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        with llg.from_case(case, form="stencil", tile=tile) as h:
            E, j = [], 0
            for n in chunks:
                E.append(h.steps(n, integ, 0.5, 0.05, kT=kT, seed=3, step0=j, energies=True))
                j += n
            m = h.get_state()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return np.concatenate(E), m


@pytest.mark.parametrize("integ", [2, 5, 6])
@pytest.mark.parametrize("spt", [1, 2])
@pytest.mark.parametrize("T", [0.0, 10.0])
@pytest.mark.parametrize("which", ["tri", "tri_ragged", "sq_bmap", "film_open_z", "sq_open"])
def test_fused_step_equals_two_stage_kernels(which, T, spt, integ):
    """Integration 2, SIB and Depondt as one kernel per step (stage 1
    recomputed on the halo in LDS; for SIB and Depondt with the ring's own
    thermal field, Depondt carrying the predictor's field), with one or two
    sites per thread, against the two stage kernels: states and energies bit
    for bit, on periodic and open lattices, lattices that are not whole
    tiles, a per-site field, eager steps (30) and graph-replayed chunks
    (64 + 7)."""
    from matjes_amd import workloads
    if which == "tri":
        s = dict(workloads.TRIANGULAR_J3DKB, size=[128, 96, 1], seed=5)
    elif which == "tri_ragged":
        s = dict(workloads.TRIANGULAR_J3DKB, size=[100, 90, 1], seed=6)
    elif which == "sq_bmap":
        s = dict(workloads.SQUARE_JDKB, size=[64, 72, 1], seed=7, bmap_sigma=0.3)
    elif which == "film_open_z":
        s = dict(workloads.FILM_DIP, size=[32, 32, 6], seed=8, dipolar=False)
    else:
        s = dict(workloads.SQUARE_JDKB, size=[80, 64, 1], seed=9, periodic=[False, True, False])
    c = workloads.build_case(s)
    kT = T * 0.000086173324
    for chunks in ((30,), (64, 7)):
        a = _run_fuse(c, False, None, kT, chunks=chunks, integ=integ)
        b = _run_fuse(c, True, None, kT, chunks=chunks, spt=spt, integ=integ)
        assert np.array_equal(a[1], b[1]), (which, chunks, spt, integ)
        assert np.array_equal(a[0], b[0]), (which, chunks, spt, integ)


@pytest.mark.parametrize("integ", [2, 5, 6])
def test_fused_last_stage_field_equals_two_stage(integ):
    """The field a fused step keeps for the host (mj_keep_last_field: the last
    stage's B_eff, thermal term excluded) against the two stage kernels' on a
    lattice above the compiled-kernel threshold (72x64), with the Philox
    field: field and state bit for bit."""
    from matjes_amd import llg, workloads
    c = workloads.build_case(dict(workloads.TRIANGULAR_J3DKB, size=[72, 64, 1], seed=5))
    out = {}
    for fuse in ("1", "0"):
        old = os.environ.get("MJ_FUSE")
        os.environ["MJ_FUSE"] = fuse
        try:
            with llg.from_case(c, form="stencil") as h:
                h.keep_last_field(True)
                h.steps(70, integ, 0.5, 0.05, kT=10.0 * 0.000086173324, seed=9)
                out[fuse] = (h.last_field(), h.get_state())
        finally:
            if old is None:
                os.environ.pop("MJ_FUSE", None)
            else:
                os.environ["MJ_FUSE"] = old
    assert np.array_equal(out["1"][0], out["0"][0]) and np.array_equal(out["1"][1], out["0"][1])
    assert np.max(np.abs(out["1"][0])) > 0.0


def _run_env(case, env, integ, steps, kT=0.0, tile=None):
    from matjes_amd import llg
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        with llg.from_case(case, form="stencil", tile=tile) as h:
            E = h.steps(steps, integ, 0.5, 0.05, kT=kT, seed=3, energies=True)
            m = h.get_state()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return E, m


@pytest.mark.parametrize("integ,T", [(2, 0.0), (2, 10.0), (1, 0.0), (5, 0.0), (6, 0.0)])
@pytest.mark.parametrize("which", ["tri", "sq_bmap", "film"])
def test_persistent_stage_kernels_equal_one_tile_per_workgroup(which, integ, T):
    """Stage 1 as persistent workgroups (next tile filled into a second LDS
    buffer while the current one is computed, MJ_PERSIST=1) against one tile
    per workgroup: states and energies bit for bit, eager (30 steps) and graph
    replayed (64 steps)."""
    from matjes_amd import workloads
    if which == "tri":
        s = dict(workloads.TRIANGULAR_J3DKB, size=[128, 96, 1], seed=1)
    elif which == "sq_bmap":
        s = dict(workloads.SQUARE_JDKB, size=[64, 72, 1], seed=2, bmap_sigma=0.3)
    else:
        s = dict(workloads.FILM_DIP, size=[32, 32, 8], seed=3, dipolar=False)
    c = workloads.build_case(s)
    kT = T * 0.000086173324
    for n in (30, 64):
        a = _run_env(c, {"MJ_PERSIST": "0"}, integ, n, kT)
        b = _run_env(c, {"MJ_PERSIST": "1"}, integ, n, kT)
        assert np.array_equal(a[1], b[1]), (which, integ, n)
        assert np.array_equal(a[0], b[0]), (which, integ, n)
