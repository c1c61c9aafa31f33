"""Differentiable SIB trajectory on the device (torch.autograd.Function).

The reference has no gradients.  This mirrors its Hamiltonian set-up (the M-M
terms of src/Hamiltonian_type/set_Hamiltonians.f90 with J, D, K scaled by
the parameters, the Zeeman "BM" term of src/Energy/zeeman.f90:30-86 with a
per-site B order parameter b_ext + h_i) and differentiates N SIB steps at
T = 0 with the hand-written adjoint of libmatjes_hip.so (mj_adjoint_*):

    m_N = SIBTrajectory.apply(theta, b_ext, hmap, m0, runner)

theta: parameter values of the unit terms named when the runner was built
(eV, e.g. (J1, D, K)); b_ext: (3,) Tesla; hmap: (N, 3) Tesla per site;
m0: (N, 3).  hmap and m0 live on the GPU (float64); theta and b_ext may live
anywhere.  The backward pass returns gradients for all four.
"""
import ctypes

import numpy as np

from . import _lib
from . import lattice as Lt
from .llg import Hamiltonian


def unit_blocks(lat, names):
    """Stencil blocks {offset: 3x3} of the unit terms (J = D = K = 1 eV)."""
    out = []
    for nm in names:
        H = Lt.Hamiltonian(lat)
        if nm.startswith("J"):
            H.add_exchange(int(nm[1:]), 1.0)
        elif nm == "D":
            H.add_dmi(1, 1.0, "Ir")
        elif nm == "K":
            H.add_anisotropy([0.0, 0.0, 1.0], 1.0)
        else:
            raise ValueError("unknown parameter %r (J1.., D, K)" % nm)
        out.append(dict(H.mm_blocks))
    return out


def _tp(t):
    return ctypes.c_void_p(t.data_ptr())


class SIBRunner:
    """Device context for the differentiable trajectory of one lattice."""

    def __init__(self, lat, names=("J1", "D", "K"), nsteps=2000, timestep=1.0, damping=0.2, ckpt_every=100,
                 device=0):
        self.lat = lat
        self.names = tuple(names)
        self.nsteps, self.timestep, self.damping, self.ckpt_every = int(nsteps), float(timestep), float(damping), \
            int(ckpt_every)
        self.n = lat.ncell
        self.h = Hamiltonian(lat.dim_lat, lat.periodic, device=device)
        Hz = Lt.Hamiltonian(lat)
        Hz.add_zeeman([0.0, 0.0, 1.0])
        self.h.add_coo(_lib.MJ_OP_B, _lib.MJ_OP_M, *Hz.coo_zeeman())
        blocks = unit_blocks(lat, self.names)
        nofs = np.array([len(b) for b in blocks], dtype=np.int32)
        ofs = np.array([o for b in blocks for o in b.keys()], dtype=np.int32).reshape(-1)
        blk = np.array([np.asarray(v, dtype=np.float64).reshape(9) for b in blocks for v in b.values()]).reshape(-1)
        self._L = _lib.load()
        _lib.check(self._L.mj_adjoint_setup(self.h.ctx, len(blocks), ctypes.c_void_p(nofs.ctypes.data),
                                            ctypes.c_void_p(ofs.ctypes.data), ctypes.c_void_p(blk.ctypes.data)),
                   "mj_adjoint_setup")
        self._last = None   # id of the inputs of the last forward

    def forward(self, theta, b_ext, hmap, m0, out):
        import torch
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64))
        be = np.ascontiguousarray(np.asarray(b_ext, dtype=np.float64))
        if th.shape != (len(self.names),) or be.shape != (3,):
            raise ValueError("theta must have %d entries and b_ext 3" % len(self.names))
        for t in (m0, out) + ((hmap,) if hmap is not None else ()):
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.shape == (self.n, 3)):
                raise ValueError("m0, hmap and the output must be contiguous float64 (N,3) GPU tensors")
        torch.cuda.current_stream().synchronize()
        _lib.check(self._L.mj_adjoint_forward(self.h.ctx, ctypes.c_void_p(th.ctypes.data),
                                              ctypes.c_void_p(be.ctypes.data),
                                              None if hmap is None else _tp(hmap), _tp(m0), self.nsteps,
                                              self.timestep, self.damping, self.ckpt_every, _tp(out)),
                   "mj_adjoint_forward")
        self._last = (th.tobytes(), be.tobytes(), None if hmap is None else hmap.data_ptr(), m0.data_ptr())
        return out

    def backward(self, gfinal, ghmap, gm0):
        import torch
        gth = np.zeros(len(self.names))
        gb = np.zeros(3)
        torch.cuda.current_stream().synchronize()
        _lib.check(self._L.mj_adjoint_backward(self.h.ctx, _tp(gfinal), ctypes.c_void_p(gth.ctypes.data),
                                               ctypes.c_void_p(gb.ctypes.data),
                                               None if ghmap is None else _tp(ghmap),
                                               None if gm0 is None else _tp(gm0)), "mj_adjoint_backward")
        return gth, gb

    def close(self):
        self.h.close()


def _np(t):
    import torch
    if isinstance(t, torch.Tensor):
        return t.detach().to("cpu", torch.float64).numpy()
    return np.asarray(t, dtype=np.float64)


def _make_function():
    import torch

    class SIBTrajectory(torch.autograd.Function):
        """m_N of N SIB steps; backward by the device adjoint with checkpoints."""

        @staticmethod
        def forward(ctx, theta, b_ext, hmap, m0, runner):
            m0c = m0.detach().contiguous()
            hc = None if hmap is None else hmap.detach().contiguous()
            out = torch.empty_like(m0c)
            runner.forward(_np(theta), _np(b_ext), hc, m0c, out)
            ctx.runner = runner
            ctx.key = runner._last
            # the tensor inputs go through save_for_backward: autograd then
            # refuses a backward after an in-place change of m0 or hmap (the
            # runner's checkpoints would belong to the old values)
            ctx.save_for_backward(hmap, m0)
            ctx.params = (_np(theta), _np(b_ext))
            ctx.meta = (theta, b_ext)
            return out

        @staticmethod
        def backward(ctx, g):
            runner = ctx.runner
            hmap, m0 = ctx.saved_tensors
            th, be = ctx.params
            m0c = m0.detach().contiguous()
            hc = None if hmap is None else hmap.detach().contiguous()
            if runner._last != ctx.key:      # another forward ran since: recompute this one
                runner.forward(th, be, hc, m0c, torch.empty_like(m0c))
            g = g.detach().contiguous()
            ghm = torch.empty_like(m0c) if ctx.needs_input_grad[2] else None
            gm0 = torch.empty_like(m0c) if ctx.needs_input_grad[3] else None
            gth, gb = runner.backward(g, ghm, gm0)
            theta, b_ext = ctx.meta

            def like(v, ref):
                if isinstance(ref, torch.Tensor):
                    return torch.as_tensor(v, dtype=ref.dtype, device=ref.device)
                return None
            return (like(gth, theta) if ctx.needs_input_grad[0] else None,
                    like(gb, b_ext) if ctx.needs_input_grad[1] else None,
                    ghm, gm0, None)

    return SIBTrajectory


_fn = None


def sib_trajectory(theta, b_ext, hmap, m0, runner):
    """Differentiable m_N (see the module docstring)."""
    global _fn
    if _fn is None:
        _fn = _make_function()
    return _fn.apply(theta, b_ext, hmap, m0, runner)
