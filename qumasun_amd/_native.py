"""ctypes binding of libqumasun_gpu.so (the C ABI declared in include/qumasun_gpu.h).

The library is built in-tree by qumasun_amd/build.py.  Loading fails loudly when it is
missing: there is no CPU fallback for any operation of this package.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libqumasun_gpu.so")

QG_OK = 0
QG_F64 = 0
QG_F32 = 1
QG_Z128 = 2
QG_C64 = 3
QG_RADIAL_GAUSSIAN = 0
QG_RADIAL_TABLE = 1
QG_PREC_F64 = 0
QG_PREC_MIXED = 1
QG_PREC_F32 = 2
QG_PRECOND_NONE = 0
QG_PRECOND_KINETIC = 1
QG_PRECOND_LEARNED = 2
QG_BASIS_REFERENCE = 0
QG_BASIS_ORTHO = 1
QG_ORTH_CHOLQR = 0
QG_ORTH_SVQB = 1
QG_KERNELS_AUTO = 0
QG_KERNELS_ROWS = 1
QG_KERNELS_TABLES = 2

c_int, c_double, c_void_p, c_int64, c_uint32, c_uint64 = (ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                                           ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64)
P = ctypes.POINTER


class ProjSpec(ctypes.Structure):
    """qg_proj_spec"""
    _fields_ = [("l", c_int), ("energy", c_double), ("cutoff", c_double), ("radial_kind", c_int),
                ("sigma", c_double), ("amplitude", c_double), ("n_table", c_int),
                ("xi_min", c_double), ("xi_delta", c_double), ("r_min", c_double),
                ("table", P(c_double))]


class VpsInfo(ctypes.Structure):
    """qg_vps_info"""
    _fields_ = [("valence_electron", c_int), ("num_projectors", c_int), ("num_radial_grids", c_int),
                ("has_pcc_charge", c_int), ("total_projector_lm", c_int), ("max_projector_l", c_int),
                ("cutoff_vlocal", c_double), ("cutoff_r", c_double * 4), ("xi_min", c_double),
                ("xi_delta", c_double), ("pcc_charge_integral", c_double)]


# void (*op)(void* user, int dtype, int nbands, const void* x, void* ax, void* stream)
OP_CALLBACK = ctypes.CFUNCTYPE(None, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p)
# void (*watch)(void* user, int sector, int64_t count)
WATCH_CALLBACK = ctypes.CFUNCTYPE(None, c_void_p, c_int, c_int64)


class LobpcgOpts(ctypes.Structure):
    """qg_lobpcg_opts"""
    _fields_ = [("precision", c_int), ("max_iter", c_int), ("tol", c_double),
                ("refresh_period", c_int), ("rollback_limit", c_double), ("precond", c_int),
                ("precond_shift", c_double), ("verbose", c_int), ("switch_tol", c_double),
                ("stall_window", c_int), ("op", OP_CALLBACK), ("op_user", c_void_p), ("basis", c_int),
                ("fallback_shift", c_double), ("watch", WATCH_CALLBACK), ("watch_user", c_void_p),
                ("orth", c_int), ("gram_chunks", c_int), ("reortho", c_double),
                ("nconv", c_int)]


class HpsiPlan(ctypes.Structure):
    """qg_hpsi_plan"""
    _fields_ = [("kernels", c_int), ("rows_ty", c_int), ("rows_bands", c_int), ("rows_tz", c_int),
                ("rows_tzp", c_int), ("rows_nbp", c_int)]


class LobpcgStats(ctypes.Structure):
    """qg_lobpcg_stats"""
    _fields_ = [("iterations", c_int), ("rollbacks", c_int), ("hpsi_calls", c_int),
                ("max_residual", c_double), ("t_hpsi", c_double), ("t_gram", c_double),
                ("t_update", c_double), ("t_eig", c_double), ("t_other", c_double),
                ("t_precond", c_double), ("refine_iterations", c_int), ("fallback_iter", c_int)]


# name -> (restype, argtypes)
SIGNATURES = {
    "qg_last_error": (ctypes.c_char_p, []),
    "qg_version": (c_int, []),
    "qg_device_count": (c_int, []),
    "qg_grid_create": (c_int, [c_int, c_int, c_int, c_double, c_double, c_double, P(c_void_p)]),
    "qg_grid_destroy": (c_int, [c_void_p]),
    "qg_laplacian8_apply": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p, c_int64,
                                    c_double, c_int, c_void_p]),
    "qg_nonlocal_create": (c_int, [c_void_p, c_int, P(c_double), P(c_int), c_int, P(c_int),
                                   P(ProjSpec), P(c_void_p)]),
    "qg_nonlocal_update": (c_int, [c_void_p, c_int, P(c_double), P(c_int)]),
    "qg_vps_load": (c_int, [ctypes.c_char_p, P(c_void_p)]),
    "qg_vps_destroy": (c_int, [c_void_p]),
    "qg_vps_get_info": (c_int, [c_void_p, P(VpsInfo)]),
    "qg_vps_get_radial": (c_int, [c_void_p, P(c_double), P(c_double), P(c_double)]),
    "qg_vps_get_projector": (c_int, [c_void_p, c_int, P(c_int), P(c_double), P(c_double), P(c_double),
                                     P(c_double)]),
    "qg_vps_projector_specs": (c_int, [c_void_p, P(ProjSpec), c_int]),
    "qg_nonlocal_destroy": (c_int, [c_void_p]),
    "qg_nonlocal_num_coef": (c_int, [c_void_p]),
    "qg_nonlocal_projector_form": (c_int, [c_void_p]),
    "qg_gram_schmidt": (c_int, [c_int, c_int64, c_int, c_void_p, c_int64, c_double, c_void_p]),
    "qg_nonlocal_project": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p, c_void_p]),
    "qg_complex_from_planar": (c_int, [c_int64, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "qg_complex_to_planar": (c_int, [c_int64, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "qg_kpoint_create": (c_int, [c_void_p, c_double, c_double, c_double, P(c_void_p)]),
    "qg_kpoint_destroy": (c_int, [c_void_p]),
    "qg_kpoint_apply": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p, c_int64, c_void_p]),
    "qg_kpoint_lobpcg": (c_int, [c_void_p, c_int, c_void_p, c_int64, P(c_double), c_int, c_double, c_int, c_double,
                                 P(c_int), P(c_double), c_void_p]),
    "qg_gradient8_apply": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64,
                                   c_double, c_void_p]),
    "qg_nonlocal_energy_force": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, P(c_double), P(c_double),
                                         P(c_double), c_void_p]),
    "qg_nonlocal_apply": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p, c_int64,
                                  c_int, c_void_p]),
    "qg_hamiltonian_create": (c_int, [c_void_p, P(c_double), c_void_p, P(c_void_p)]),
    "qg_hamiltonian_destroy": (c_int, [c_void_p]),
    "qg_hamiltonian_set_potential": (c_int, [c_void_p, P(c_double)]),
    "qg_hamiltonian_set_potential_device": (c_int, [c_void_p, c_void_p, c_void_p]),
    "qg_occupancy_fermi": (c_int, [c_int, P(c_double), c_double, c_double, P(c_double), P(c_double)]),
    "qg_density": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, P(c_double), c_void_p, c_void_p]),
    "qg_density_mix": (c_int, [c_void_p, c_void_p, c_void_p, c_double, P(c_double), P(c_double), c_void_p]),
    "qg_hartree_potential": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "qg_xc_lda": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "qg_scf_potential": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p]),
    "qg_hamiltonian_apply": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p, c_int64,
                                     c_void_p]),
    "qg_hamiltonian_apply_noalloc": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p, c_int64,
                                             c_void_p]),
    "qg_default_hpsi_plan": (None, [P(HpsiPlan)]),
    "qg_hamiltonian_set_plan": (c_int, [c_void_p, P(HpsiPlan)]),
    "qg_hamiltonian_set_timing": (c_int, [c_void_p, c_int]),
    "qg_hamiltonian_last_times": (c_int, [c_void_p, P(c_double), P(c_double)]),
    "qg_hamiltonian_last_chunks": (c_int, [c_void_p, P(c_int)]),
    "qg_lobpcg_default_opts": (None, [P(LobpcgOpts)]),
    "qg_lobpcg_d_multi": (c_int, [c_void_p, P(LobpcgOpts), c_int, P(c_double), c_void_p, c_void_p,
                                  c_int, P(LobpcgStats)]),
    "qg_precond_apply": (c_int, [c_void_p, c_int, c_double, c_int, c_int, c_void_p, c_int64, c_void_p]),
    "qg_gram": (c_int, [c_int, c_void_p, c_void_p, c_int64, c_int, c_int, c_int64, c_int64, c_void_p, c_int,
                        c_void_p]),
    "qg_gram_bf16": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int, c_int64, c_int64, c_void_p, c_int,
                             c_int, c_void_p]),
    "qg_cnn_num_params": (c_int, []),
    "qg_cnn_create": (c_int, [c_void_p, c_double, c_int, c_uint64, P(c_void_p)]),
    "qg_cnn_destroy": (c_int, [c_void_p]),
    "qg_cnn_get_params": (c_int, [c_void_p, P(ctypes.c_float), c_int]),
    "qg_cnn_set_params": (c_int, [c_void_p, P(ctypes.c_float), c_int]),
    "qg_cnn_network": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "qg_cnn_loss_grad": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_double, P(c_double),
                                 P(ctypes.c_float)]),
    "qg_cnn_collect": (c_int, [c_void_p, c_int, c_void_p, P(c_double), c_double, c_double, c_int, P(c_int)]),
    "qg_cnn_num_pairs": (c_int, [c_void_p]),
    "qg_cnn_train": (c_int, [c_void_p, c_int, c_int, c_double, c_uint64, P(c_double), P(c_double)]),
    "qg_cnn_apply": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p]),
    "qg_hamiltonian_set_cnn": (c_int, [c_void_p, c_void_p]),
    "qg_batch_hamiltonian_apply": (c_int, [c_int, P(c_void_p), c_int, P(c_int), P(c_void_p), P(c_void_p),
                                           c_void_p]),
    "qg_autotune_hamiltonian": (c_int, [c_void_p, c_int, c_int, c_int, P(c_int), P(c_int), P(c_double)]),
    "qg_batch_lobpcg": (c_int, [c_int, P(c_void_p), P(LobpcgOpts), P(c_int), P(P(c_double)), P(c_void_p),
                                P(c_void_p), c_int, P(LobpcgStats), c_int]),
    "qg_synth_diamond_atoms": (c_int, [c_int, c_double, P(c_double)]),
    "qg_synth_vloc": (c_int, [c_int, c_int, c_int, c_double, c_double, c_double, c_int, P(c_double),
                              c_double, c_double, c_double, c_uint32, P(c_double)]),
    "qg_synth_random_states": (c_int, [c_int64, c_int, c_double, c_uint32, P(c_double)]),
}

_lib = None


class QGError(RuntimeError):
    pass


def lib():
    """Load libqumasun_gpu.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QGError(f"{LIB_PATH} is missing: build it with `python -m qumasun_amd.build` "
                          "(there is no CPU fallback)")
        # torch's wheel ships its own HIP and HSA runtimes under the same sonames as
        # /opt/rocm's.  Loaded after torch, this library's libamdhip64.so.7 and
        # libhsa-runtime64.so.1 resolve to torch's copies (one runtime in the process); loaded
        # before it, torch later maps a second HIP and HSA runtime onto the same device, and its
        # first device call fails (hipErrorNoDevice) or crashes.  So torch comes first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(rc):
    if rc != QG_OK:
        msg = lib().qg_last_error().decode(errors="replace")
        raise QGError(f"qumasun_gpu error {rc}: {msg}")
    return rc


def dptr(a):
    """ctypes double* of a C-contiguous float64 numpy array."""
    return a.ctypes.data_as(P(c_double))
