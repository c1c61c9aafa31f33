"""ctypes binding of libmatjes_hip.so (the C-ABI in include/matjes_hip.h).

The device library is the only compute path: if it is missing or fails to
load, every entry point raises instead of falling back to anything.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmatjes_hip.so")

MJ_OK, MJ_EINVAL, MJ_EHIP, MJ_ESTATE, MJ_ENOMEM, MJ_ENOTSUP = 0, -1, -2, -3, -4, -5
MJ_MC_SPHERE, MJ_MC_ISING = 0, 1
MJ_EULER, MJ_HEUN, MJ_RK, MJ_SIB, MJ_DEPONDT = 1, 2, 4, 5, 6
MJ_OP_M, MJ_OP_B = 1, 3

# every symbol declared in include/matjes_hip.h
EXPORTS = [
    "mj_last_error", "mj_create", "mj_destroy", "mj_add_hamiltonian_coo", "mj_add_hamiltonian_stencil",
    "mj_set_bfield", "mj_set_fft_operator", "mj_set_state", "mj_get_state", "mj_set_state_device",
    "mj_get_state_device", "mj_eff_field", "mj_energy", "mj_energy_terms", "mj_steps", "mj_steps_noise", "mj_profile_steps",
    "mj_steps_adaptive",
    "mj_thermal_normals", "mj_set_tile", "mj_adjoint_setup", "mj_adjoint_forward", "mj_adjoint_backward",
    "mj_batch_create", "mj_batch_destroy", "mj_batch_add", "mj_batch_set_pt", "mj_batch_set_tiles",
    "mj_batch_set_launch",
    "mj_batch_steps", "mj_batch_profile", "mj_batch_get_state", "mj_batch_energies", "mj_batch_pt_state",
    "mj_batch_info", "mj_energy_single", "mj_mc_sweeps", "mj_mc_set_cone", "mj_mc_stats", "mj_term_mult",
    "mj_mc_pt_setup", "mj_mc_pt_sweeps", "mj_mc_pt_state", "mj_mc_pt_get_state", "mj_set_bfield_maps",
    "mj_steps_driven", "mj_energy_noalloc", "mj_set_rk_order", "mj_math_probe", "mj_keep_last_field",
    "mj_last_field", "mj_steps_temperature",
]

_lib = None


class MatjesError(RuntimeError):
    pass


def _sig(L):
    c_int, c_i64, c_u64, c_dbl = ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
    vp = ctypes.c_void_p
    L.mj_last_error.argtypes = [ctypes.c_char_p, c_int]
    L.mj_create.argtypes = [ctypes.POINTER(vp), c_int, vp, vp, c_int]
    L.mj_destroy.argtypes = [ctypes.POINTER(vp)]
    L.mj_add_hamiltonian_coo.argtypes = [vp, c_int, c_int, c_i64, vp, vp, vp]
    L.mj_add_hamiltonian_stencil.argtypes = [vp, c_int, vp, vp]
    L.mj_set_bfield.argtypes = [vp, vp, c_i64]
    L.mj_set_fft_operator.argtypes = [vp, c_int, vp, vp, vp]
    for f in ("mj_set_state", "mj_get_state", "mj_set_state_device", "mj_get_state_device", "mj_eff_field",
              "mj_last_field"):
        getattr(L, f).argtypes = [vp, vp, c_i64]
    L.mj_keep_last_field.argtypes = [vp, c_int]
    L.mj_energy.argtypes = [vp, vp]
    L.mj_energy_noalloc.argtypes = [vp, vp]
    L.mj_energy_terms.argtypes = [vp, vp, vp]
    L.mj_energy_single.argtypes = [vp, c_i64, vp, vp, vp]
    L.mj_term_mult.argtypes = [vp, c_int, c_int, vp, c_i64, vp, c_i64, c_dbl, c_dbl]
    L.mj_mc_pt_setup.argtypes = [vp, c_int, vp, c_dbl]
    L.mj_set_bfield_maps.argtypes = [vp, c_int, vp, c_i64]
    L.mj_steps_driven.argtypes = [vp, c_int, c_int, c_dbl, c_dbl, c_dbl, c_u64, c_u64, vp, vp, vp]
    L.mj_mc_pt_sweeps.argtypes = [vp, c_int, c_int, c_u64, c_u64, c_int, c_int, vp, vp]
    L.mj_mc_pt_state.argtypes = [vp, vp, vp, vp, vp]
    L.mj_mc_pt_get_state.argtypes = [vp, c_int, vp, c_i64]
    L.mj_mc_sweeps.argtypes = [vp, c_int, c_int, c_dbl, c_u64, c_u64, c_int, vp, vp]
    L.mj_mc_set_cone.argtypes = [vp, c_dbl]
    L.mj_mc_stats.argtypes = [vp, vp, vp]
    L.mj_steps.argtypes = [vp, c_int, c_int, c_dbl, c_dbl, c_dbl, c_u64, c_u64, vp]
    L.mj_steps_noise.argtypes = [vp, c_int, c_int, c_dbl, c_dbl, c_dbl, vp, vp]
    L.mj_steps_adaptive.argtypes = [vp, c_int, c_int, vp, c_dbl, c_dbl, c_u64, c_u64, c_dbl, vp, vp]
    L.mj_steps_temperature.argtypes = [vp, c_int, c_int, c_dbl, c_dbl, vp, c_u64, c_u64, vp]
    L.mj_profile_steps.argtypes = [vp, c_int, c_int, c_dbl, c_dbl, c_dbl, c_u64, c_int, vp, vp, vp]
    L.mj_thermal_normals.argtypes = [vp, c_u64, c_u64, vp, c_i64]
    L.mj_set_tile.argtypes = [vp, c_int, c_int, c_int]
    L.mj_set_rk_order.argtypes = [vp, c_int]
    L.mj_math_probe.argtypes = [vp, c_int, c_i64, vp, vp]
    L.mj_adjoint_setup.argtypes = [vp, c_int, vp, vp, vp]
    L.mj_adjoint_forward.argtypes = [vp, vp, vp, vp, vp, c_int, c_dbl, c_dbl, c_int, vp]
    L.mj_adjoint_backward.argtypes = [vp, vp, vp, vp, vp, vp]
    L.mj_batch_create.argtypes = [ctypes.POINTER(vp), c_int]
    L.mj_batch_destroy.argtypes = [ctypes.POINTER(vp)]
    L.mj_batch_add.argtypes = [vp, vp, c_dbl, c_dbl, c_u64, vp]
    L.mj_batch_set_pt.argtypes = [vp, c_int, c_int, vp, c_int, c_u64]
    L.mj_batch_set_tiles.argtypes = [vp, vp]
    L.mj_batch_set_launch.argtypes = [vp, ctypes.c_int, ctypes.c_int64]
    L.mj_batch_steps.argtypes = [vp, c_int, c_int, c_dbl]
    L.mj_batch_profile.argtypes = [vp, c_int, c_int, c_dbl, vp]
    L.mj_batch_get_state.argtypes = [vp, c_int, vp, c_i64]
    L.mj_batch_energies.argtypes = [vp, vp]
    L.mj_batch_pt_state.argtypes = [vp, vp, vp, vp]
    L.mj_batch_info.argtypes = [vp, vp, vp, vp]
    for f in EXPORTS:
        getattr(L, f).restype = c_int


def load():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise MatjesError("libmatjes_hip.so not built (run __graft_entry__.build() or "
                              "python -m matjes_amd.build); there is no CPU fallback")
        _torch_runtime_first()
        L = ctypes.CDLL(LIB_PATH)
        _sig(L)
        _lib = L
    return _lib


def _torch_runtime_first():
    """PyTorch ships its own HIP/HSA runtime next to the system one this library
    links.  Both can live in one process only if PyTorch's starts first (the
    other order leaves PyTorch without devices), so when PyTorch is installed
    and sees a GPU, start it before the library makes its first HIP call."""
    try:
        import torch
    except ImportError:
        return
    try:
        if torch.cuda.device_count() > 0:
            torch.cuda.init()
    except RuntimeError:
        pass


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().mj_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc, what):
    if rc != MJ_OK:
        raise MatjesError("%s failed (%d): %s" % (what, rc, last_error()))
