"""mj_energy_noalloc (static scratch, no device allocation) against mj_energy,
through the C-ABI: two calls in a row on two contexts sharing the GPU's
scratch array, the second on a smaller lattice, each equal to mj_energy of
the same state (the same partials in the same order: bit for bit)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _energy_noalloc(h):
    import torch
    from matjes_amd import _lib
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    _lib.check(_lib.load().mj_energy_noalloc(h.ctx, ctypes.c_void_p(out.data_ptr())), "mj_energy_noalloc")
    torch.cuda.synchronize()
    return float(out.cpu()[0])


def _case(size, seed):
    from matjes_amd import workloads
    return workloads.build_case(dict(workloads.TRIANGULAR_J3DKB, size=size, seed=seed))


def test_noalloc_energy_equals_mj_energy_on_two_calls():
    from matjes_amd import llg
    from matjes_amd import _lib
    big, small = _case([128, 96, 1], 41), _case([48, 32, 1], 42)
    with llg.from_case(big) as hb, llg.from_case(small) as hs:
        Eb = hb.energy()            # builds the device Hamiltonian (allocates)
        Es = hs.energy()
        got_b = _energy_noalloc(hb)
        got_s = _energy_noalloc(hs)   # second call, smaller lattice, same scratch array
        assert got_b == Eb, (got_b, Eb)
        assert got_s == Es, (got_s, Es)
        # after steps the state changed: the scratch array holds the last call's
        # partials, which the next call overwrites
        hs.steps(5, _lib.MJ_HEUN, 0.5, 0.05)
        assert _energy_noalloc(hs) == hs.energy()
        assert np.isfinite(got_b) and np.isfinite(got_s)


def test_noalloc_energy_refuses_a_hamiltonian_not_on_the_device():
    from matjes_amd import llg
    import torch
    from matjes_amd import _lib
    c = _case([16, 16, 1], 43)
    with llg.from_case(c) as h:   # no energy / steps call yet: the device Hamiltonian is not built
        out = torch.zeros(1, dtype=torch.float64, device="cuda:0")
        rc = _lib.load().mj_energy_noalloc(h.ctx, ctypes.c_void_p(out.data_ptr()))
        assert rc == _lib.MJ_ESTATE
