"""Device adjoint of the SIB trajectory against the FP64 reference gradient (MI355X).

The reference gradient is the complex-step derivative of the oracle's SIB map
(oracle/adjoint_ref.py, exact to rounding); central differences (h = 1e-6)
of the device forward and a single-spin closed form are checked as well.
Gradients agree to 1e-9 relative (the map and its transpose are summed in a
different order on the device); checkpoint spacing changes nothing bitwise."""
import numpy as np
import pytest

from matjes_amd import lattice as L
from matjes_amd import workloads

pytestmark = pytest.mark.gpu

NAMES = ["J1", "D", "K"]
THETA = np.array([1e-3, 0.3e-3, -0.05e-3])


def _lat(size, periodic=(True, True, False), kind=None):
    s = dict(kind or workloads.SQUARE_JDKB, size=list(size), periodic=list(periodic))
    return workloads.make_lattice(s)


def _inputs(lat, seed=5, hsig=0.3):
    import torch
    rng = np.random.default_rng(seed)
    h = hsig * rng.normal(size=(lat.ncell, 3))
    m0 = workloads.random_sphere(lat.ncell, seed)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    return h, m0, dev(h), dev(m0)


def _run(lat, names, theta, bext, h, m0, nsteps, dt, alpha, ckpt, tgt):
    import torch
    from matjes_amd import adjoint
    r = adjoint.SIBRunner(lat, names, nsteps=nsteps, timestep=dt, damping=alpha, ckpt_every=ckpt)
    try:
        out = torch.empty_like(m0)
        r.forward(theta, bext, h, m0, out)
        g = 2 * (out - torch.tensor(tgt, dtype=torch.float64, device="cuda"))
        ghm, gm0 = torch.empty_like(m0), torch.empty_like(m0)
        gth, gb = r.backward(g.contiguous(), ghm, gm0)
        torch.cuda.synchronize()
        return out.cpu().numpy(), gth, gb, ghm.cpu().numpy(), gm0.cpu().numpy()
    finally:
        r.close()


def test_forward_matches_plain_sib_and_oracle(oracle_lib):
    import torch
    from matjes_amd import llg
    lat = _lat((12, 10, 1))
    h, m0, hd, m0d = _inputs(lat)
    bext = np.array([0.0, 0.0, 0.5])
    tgt = np.zeros_like(m0)
    mN, *_ = _run(lat, NAMES, THETA, bext, hd, m0d, 50, 1.0, 0.2, 10, tgt)
    H = L.Hamiltonian(lat)
    H.add_exchange(1, THETA[0])
    H.add_dmi(1, THETA[1], "Ir")
    H.add_anisotropy([0.0, 0.0, 1.0], THETA[2])
    H.add_zeeman(bext[None, :] + h)
    mo = oracle_lib.Oracle(H).steps(m0, 50, 5, 1.0, 0.2)
    assert np.max(np.abs(mN - mo)) < 1e-13
    with llg.Hamiltonian(lat.dim_lat, lat.periodic) as d:
        d.load_terms(H, form="stencil")
        d.set_state(m0)
        d.steps(50, llg.SIB, 1.0, 0.2)
        mp = d.get_state()
    assert np.max(np.abs(mN - mp)) < 1e-13


@pytest.mark.parametrize("geom", ["square_periodic", "square_open", "triangular_j3"])
def test_gradients_match_complex_step_reference(geom):
    from oracle import adjoint_ref as R
    if geom == "triangular_j3":
        lat = _lat((7, 6, 1), kind=workloads.TRIANGULAR_J3DKB)
        names = ["J1", "J2", "J3", "D", "K"]
        theta = np.array([1e-3, -0.1e-3, -0.05e-3, 0.4e-3, -0.1e-3])
    else:
        lat = _lat((8, 6, 1), periodic=(True, True, False) if geom == "square_periodic" else (False, False, False))
        names, theta = NAMES, THETA
    h, m0, hd, m0d = _inputs(lat)
    bext = np.array([0.1, -0.2, 0.5])
    tgt = R.neel_skyrmion(lat.dim_lat[0], lat.dim_lat[1], 1.5)
    N, dt, al = 60, 1.0, 0.2
    _, gth, gb, ghm, gm0 = _run(lat, names, theta, bext, hd, m0d, N, dt, al, 7, tgt)
    rng = np.random.default_rng(3)
    dh, dm = rng.normal(size=h.shape), rng.normal(size=m0.shape)
    smap = R.SIBMap(lat, names)
    rth, rb, dirs = R.reference_gradients(smap, theta, bext, h, m0, tgt, N, dt, al,
                                          directions=[(dh, 0 * dm), (0 * dh, dm)])
    assert np.max(np.abs(gth - rth)) <= 1e-9 * np.max(np.abs(rth))
    assert np.max(np.abs(gb - rb)) <= 1e-9 * np.max(np.abs(rb))
    assert abs(np.sum(ghm * dh) - dirs[0]) <= 1e-9 * abs(dirs[0])
    assert abs(np.sum(gm0 * dm) - dirs[1]) <= 1e-9 * abs(dirs[1])
    # the gradient of the uniform field is the sum of the per-site one
    assert np.max(np.abs(ghm.sum(0) - gb)) <= 1e-12 * np.max(np.abs(gb))


def test_checkpoint_spacing_is_bitwise_invariant():
    lat = _lat((10, 8, 1))
    h, m0, hd, m0d = _inputs(lat)
    from oracle import adjoint_ref as R
    tgt = R.neel_skyrmion(10, 8, 2.0)
    res = [_run(lat, NAMES, THETA, [0.0, 0.0, 0.5], hd, m0d, 45, 1.0, 0.2, ck, tgt) for ck in (1, 7, 45, 100)]
    for r in res[1:]:
        for a, b in zip(res[0], r):
            assert np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("geom", ["square_periodic", "square_open", "triangular_j3", "square_128"])
def test_fused_forward_step_is_bitwise_the_two_kernels(geom):
    """The forward SIB step as one kernel (the midpoint on the tile's ring
    from twice the reach in LDS) against the two stage kernels
    (MJ_ADJ_FUSE=0): final state, midpoints through the reverse sweep, and
    every gradient bit for bit; periodic, open, reach 3 and the 128x128
    lattice of the adjoint form."""
    import os
    from oracle import adjoint_ref as R
    if geom == "triangular_j3":
        lat = _lat((9, 8, 1), kind=workloads.TRIANGULAR_J3DKB)
        names = ["J1", "J2", "J3", "D", "K"]
        theta = np.array([1e-3, -0.1e-3, -0.05e-3, 0.4e-3, -0.1e-3])
    elif geom == "square_128":
        lat = _lat((128, 128, 1))
        names, theta = NAMES, THETA
    else:
        lat = _lat((20, 14, 1), periodic=(True, True, False) if geom == "square_periodic" else (False, False, False))
        names, theta = NAMES, THETA
    h, m0, hd, m0d = _inputs(lat)
    tgt = R.neel_skyrmion(lat.dim_lat[0], lat.dim_lat[1], 3.0)
    res = {}
    old = os.environ.get("MJ_ADJ_FUSE")
    try:
        for v in ("1", "0"):
            os.environ["MJ_ADJ_FUSE"] = v
            res[v] = _run(lat, names, theta, [0.1, -0.2, 0.5], hd, m0d, 40, 1.0, 0.2, 10, tgt)
    finally:
        if old is None:
            os.environ.pop("MJ_ADJ_FUSE", None)
        else:
            os.environ["MJ_ADJ_FUSE"] = old
    for a, b in zip(res["1"], res["0"]):
        assert np.array_equal(np.asarray(a), np.asarray(b)), geom


def test_central_differences_of_device_forward():
    import torch
    from matjes_amd import adjoint
    from oracle import adjoint_ref as R
    lat = _lat((16, 16, 1))
    h, m0, hd, m0d = _inputs(lat, hsig=0.05e-3 / (2.0 * L.MU_B))
    tgt = torch.tensor(R.neel_skyrmion(16, 16, 4.0), dtype=torch.float64, device="cuda")
    bext = np.array([0.0, 0.0, 0.5])
    r = adjoint.SIBRunner(lat, NAMES, nsteps=200, timestep=1.0, damping=0.2, ckpt_every=50)
    try:
        out = torch.empty_like(m0d)

        def loss(th, b):
            r.forward(th, b, hd, m0d, out)
            return float(((out - tgt) ** 2).sum())
        loss(THETA, bext)
        gth, gb = r.backward((2 * (out - tgt)).contiguous(), None, None)
        for k in range(3):
            e = 1e-6 * abs(THETA[k])
            d = np.eye(3)[k] * e
            fd = (loss(THETA + d, bext) - loss(THETA - d, bext)) / (2 * e)
            assert abs(fd - gth[k]) <= 3e-5 * np.max(np.abs(gth))
        e = 1e-6
        fd = (loss(THETA, bext + [0, 0, e]) - loss(THETA, bext - [0, 0, e])) / (2 * e)
        assert abs(fd - gb[2]) <= 3e-5 * max(abs(gb[2]), np.max(np.abs(gth)) * abs(THETA[0]))
    finally:
        r.close()


def test_single_spin_closed_form():
    import torch
    from matjes_amd import adjoint
    lat = L.Lattice(np.eye(3), (1.0, 1.0, 1.0), (1, 1, 1), (False, False, False), [("Fe", (0, 0, 0), 2.0)])
    N, dt, b = 40, 1.0, 3.0
    r = adjoint.SIBRunner(lat, (), nsteps=N, timestep=dt, damping=0.0, ckpt_every=9)
    try:
        m0 = torch.tensor([[1.0, 0.0, 0.0]], dtype=torch.float64, device="cuda")
        out = torch.empty_like(m0)
        r.forward(np.zeros(0), [0.0, 0.0, b], None, m0, out)
        kap = dt * 2.0 * L.MU_B * b / (2 * L.HBAR)
        phi = 2 * np.arctan(kap)
        np.testing.assert_allclose(out.cpu().numpy()[0], [np.cos(N * phi), np.sin(N * phi), 0.0], atol=1e-13)
        g = torch.tensor([[0.0, 1.0, 0.0]], dtype=torch.float64, device="cuda")
        _, gb = r.backward(g, None, None)
        want = N * np.cos(N * phi) * 2 * kap / ((1 + kap * kap) * b)
        assert abs(gb[2] - want) <= 1e-12 * abs(want)
    finally:
        r.close()


def test_autograd_function():
    import torch
    from matjes_amd import adjoint
    from oracle import adjoint_ref as R
    lat = _lat((8, 6, 1))
    h, m0, hd, m0d = _inputs(lat)
    tgt = R.neel_skyrmion(8, 6, 1.5)
    r = adjoint.SIBRunner(lat, NAMES, nsteps=30, timestep=1.0, damping=0.2, ckpt_every=10)
    try:
        theta = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
        bext = torch.tensor([0.1, -0.2, 0.5], dtype=torch.float64, requires_grad=True)
        hmap = hd.clone().requires_grad_(True)
        m0t = m0d.clone().requires_grad_(True)
        mN = adjoint.sib_trajectory(theta, bext, hmap, m0t, r)
        loss = ((mN - torch.tensor(tgt, dtype=torch.float64, device="cuda")) ** 2).sum()
        loss.backward()
        rth, rb, dirs = R.reference_gradients(R.SIBMap(lat, NAMES), THETA, [0.1, -0.2, 0.5], h, m0, tgt, 30, 1.0,
                                              0.2, directions=[(np.ones_like(h), 0 * h)])
        assert np.max(np.abs(theta.grad.numpy() - rth)) <= 1e-9 * np.max(np.abs(rth))
        assert np.max(np.abs(bext.grad.numpy() - rb)) <= 1e-9 * np.max(np.abs(rb))
        assert abs(hmap.grad.sum().item() - dirs[0]) <= 1e-9 * abs(dirs[0])
        assert m0t.grad is not None and torch.isfinite(m0t.grad).all()
        # a second forward in between: backward recomputes its own trajectory
        mN2 = adjoint.sib_trajectory(theta.detach() * 1.1, bext, hmap, m0t, r)
        theta.grad = None
        mN3 = adjoint.sib_trajectory(theta, bext, hmap, m0t, r)
        adjoint.sib_trajectory(theta.detach() * 0.9, bext.detach(), hmap.detach(), m0t.detach(), r)
        ((mN3 - torch.tensor(tgt, dtype=torch.float64, device="cuda")) ** 2).sum().backward()
        assert np.max(np.abs(theta.grad.numpy() - rth)) <= 1e-9 * np.max(np.abs(rth))
        del mN2
        # an in-place change of an input between forward and backward: autograd
        # refuses the backward instead of using the old trajectory's checkpoints
        for which in ("m0", "hmap"):
            mN4 = adjoint.sib_trajectory(theta, bext, hmap, m0t, r)
            with torch.no_grad():
                (m0t if which == "m0" else hmap).mul_(1.0)
            with pytest.raises(RuntimeError, match="inplace"):
                mN4.sum().backward()
    finally:
        r.close()


def test_bad_arguments():
    from matjes_amd import _lib
    lat = _lat((4, 4, 1))
    from matjes_amd import llg
    with llg.Hamiltonian(lat.dim_lat, lat.periodic) as d:
        L_ = _lib.load()
        rc = L_.mj_adjoint_setup(d.ctx, 0, None, None, None)
        assert rc == _lib.MJ_ESTATE          # no Zeeman term registered
        rc = L_.mj_adjoint_backward(d.ctx, None, None, None, None, None)
        assert rc == _lib.MJ_EINVAL
