"""Batched systems and parallel tempering on the device (mj_batch_* of the C-ABI).

The reference runs one lattice per process and has no replica exchange; the
per-system arithmetic is that of `llg.Hamiltonian.steps` with integration 1,
2, 5 or 6 (src/dynamics/spindynamics.f90:260-323), and a batched system's
trajectory equals the single-system one bit for bit.
"""
import ctypes

import numpy as np

from . import _lib
from .llg import HEUN


STAGES, FUSED, PERSISTENT, PERSISTENT_FUSED = 0, 1, 2, 3   # MJ_BATCH_* of the C-ABI


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class Batch:
    """Many lattices advanced by one launch per integrator stage."""

    def __init__(self, device=0):
        self._L = _lib.load()
        b = ctypes.c_void_p()
        _lib.check(self._L.mj_batch_create(ctypes.byref(b), int(device)), "mj_batch_create")
        self._b = b
        self.sizes = []
        self.nstep = 0          # steps taken (segments start at multiples of the segment length)
        self.launch = FUSED

    def close(self):
        if getattr(self, "_b", None) is not None and self._b.value:
            self._L.mj_batch_destroy(ctypes.byref(self._b))
        self._b = None

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
    def b(self):
        if self._b is None or not self._b.value:
            raise _lib.MatjesError("batch destroyed")
        return self._b

    def add(self, h, damping, kT=0.0, seed=0):
        """Copy the Hamiltonian and current state of an llg.Hamiltonian (kT in eV)."""
        idx = ctypes.c_int()
        _lib.check(self._L.mj_batch_add(self.b, h.ctx, float(damping), float(kT), int(seed), ctypes.byref(idx)),
                   "mj_batch_add")
        self.sizes.append(h.ncell)
        return idx.value

    def set_pt(self, first, count, kT_ladder, every, seed=0):
        lad = np.ascontiguousarray(kT_ladder, dtype=np.float64)
        if lad.shape != (count,):
            raise ValueError("one temperature per replica")
        _lib.check(self._L.mj_batch_set_pt(self.b, int(first), int(count), _p(lad), int(every), int(seed)),
                   "mj_batch_set_pt")

    def set_tiles(self, tiles):
        t = np.ascontiguousarray(np.asarray(tiles, dtype=np.int32).reshape(-1))
        if t.size != 2 * len(self.sizes):
            raise ValueError("one (bx, by) per system")
        _lib.check(self._L.mj_batch_set_tiles(self.b, _p(t)), "mj_batch_set_tiles")

    def set_launch(self, mode=FUSED, spin=-1):
        """How integration 2's captured segments run: STAGES (one launch per
        stage), FUSED (one per step, both stages per tile; the default),
        PERSISTENT or PERSISTENT_FUSED (one launch per segment, grid barriers
        between stages or steps).  The trajectories are the same bit for bit.
        A persistent barrier gives up after `spin` sleeps of about 0.25 us
        (negative: the default, about 1 s): the run then raises and the states
        are undefined."""
        _lib.check(self._L.mj_batch_set_launch(self.b, int(mode), int(spin)), "mj_batch_set_launch")
        self.launch = int(mode)

    def steps(self, nsteps, timestep, integrator=HEUN):
        _lib.check(self._L.mj_batch_steps(self.b, int(nsteps), int(integrator), float(timestep)), "mj_batch_steps")
        self.nstep += int(nsteps)

    def profile(self, nsteps, timestep, integrator=HEUN):
        ms = ctypes.c_double()
        _lib.check(self._L.mj_batch_profile(self.b, int(nsteps), int(integrator), float(timestep), ctypes.byref(ms)),
                   "mj_batch_profile")
        self.nstep += int(nsteps)
        return ms.value

    def state(self, i):
        out = np.empty(3 * self.sizes[i])
        _lib.check(self._L.mj_batch_get_state(self.b, int(i), _p(out), out.size), "mj_batch_get_state")
        return out.reshape(-1, 3)

    def energies(self):
        E = np.empty(len(self.sizes))
        _lib.check(self._L.mj_batch_energies(self.b, _p(E)), "mj_batch_energies")
        return E

    def pt_state(self, count):
        rung = np.empty(count, dtype=np.int32)
        acc = np.empty(count, dtype=np.uint64)
        att = np.empty(count, dtype=np.uint64)
        _lib.check(self._L.mj_batch_pt_state(self.b, _p(rung), _p(acc), _p(att)), "mj_batch_pt_state")
        return rung, acc, att

    def info(self):
        sites = ctypes.c_int64()
        wg = ctypes.c_int()
        tiles = np.empty(2 * len(self.sizes), dtype=np.int32)
        _lib.check(self._L.mj_batch_info(self.b, ctypes.byref(sites), ctypes.byref(wg), _p(tiles)), "mj_batch_info")
        return sites.value, wg.value, tiles.reshape(-1, 2)
