"""germaine_amd — MI355X-native solve path of GERMAINE's FDFD modelling/FWI.

The product is ``libgfdfd.so`` (built in-tree from ``germaine_amd/csrc``):
HIP kernels for gfx950 behind two C ABIs, the six ``umfpack_zi_*`` entry
points GERMAINE links against (``include/umfpack.h``) and the ``gfdfd_*``
extension (``include/gfdfd.h``).  This module loads the library with ctypes
and declares every entry point; ``germaine_amd.fdfd`` holds the host-side
workflows (frequency batches, shot panels, gradients) that drive it.

There is no CPU fallback: device entry points fail loudly when no GPU or no
built library is present.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgfdfd.so")
_lib = None

GFDFD_OK, GFDFD_WARN_SINGULAR = 0, 1
AC, SH, TE = 1, 2, 3
FP64, FP32 = 64, 32
UMFPACK_A, UMFPACK_At, UMFPACK_Aat = 0, 1, 2


class GfdfdError(RuntimeError):
    pass


class ShotTaperC(C.Structure):
    """gfdfd_shot_taper (include/gfdfd.h)."""
    _fields_ = [("mode", C.c_int), ("nx", C.c_int), ("dh", C.c_float), ("srtradius", C.c_float),
                ("filtsize", C.c_int), ("d_srcx", C.c_void_p), ("d_srcy", C.c_void_p), ("nsrc", C.c_int),
                ("d_src_i", C.c_void_p), ("d_src_j", C.c_void_p), ("nrec", C.c_int), ("d_rec_i", C.c_void_p),
                ("d_rec_j", C.c_void_p)]


class PlanInfo(C.Structure):
    _fields_ = [("n", C.c_longlong), ("nnz", C.c_longlong), ("fronts", C.c_int), ("levels", C.c_int),
                ("max_front", C.c_int), ("max_pivots", C.c_int), ("arena_elems", C.c_longlong),
                ("work_rows", C.c_longlong), ("factor_flops", C.c_double), ("nnz_lu", C.c_double),
                ("plan_device_bytes", C.c_double)]


class FreqParams(C.Structure):
    _fields_ = [("omega2", C.c_float * 2), ("omega", C.c_float * 2), ("prof", C.POINTER(C.c_float))]


_P, _I, _D, _F, _V = C.c_void_p, C.c_int, C.c_double, C.c_float, None
_IP, _FP, _DP, _LLP = C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_longlong)

# name: (restype, argtypes) for every exported entry point of include/gfdfd.h and include/umfpack.h
SIGNATURES = {
    "gfdfd_last_error": (C.c_char_p, []),
    "gfdfd_plan_grid": (_I, [_I, _I, C.POINTER(_P)]),
    "gfdfd_plan_grid_opt": (_I, [_I, _I, _I, _I, C.POINTER(_P)]),
    "gfdfd_plan_free": (_V, [C.POINTER(_P)]),
    "gfdfd_plan_query": (_I, [_P, C.POINTER(PlanInfo)]),
    "gfdfd_plan_pattern": (_I, [_P, _P, _P]),
    "gfdfd_plan_perm": (_I, [_P, _P]),
    "gfdfd_plan_fronts": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "gfdfd_plan_index": (_I, [_P, _P]),
    "gfdfd_fd_weights": (_V, [_F, _P]),
    "gfdfd_mu0": (_F, []),
    "gfdfd_pml_profiles": (_V, [_I, _I, _I, _F, _F, _F, _F, _I, _P]),
    "gfdfd_omega": (_V, [_I, _F, _F, _P, _P]),
    "gfdfd_omega2_grad_ac": (_V, [_F, _F, _P]),
    "gfdfd_source_amp": (_V, [_I, _F, _F, _F, _P]),
    "gfdfd_assemble": (_I, [_P, _I, _F, _I, C.POINTER(FreqParams), _P, _P, _P, _P]),
    "gfdfd_factor": (_I, [_P, _I, _P, _I, _P, C.POINTER(_P)]),
    "gfdfd_refactor": (_I, [_P, _P, _P]),
    "gfdfd_refactor_graph": (_I, [_P, _P, _P]),
    "gfdfd_solve_tuning": (_I, [_I, _I, _I]),
    "gfdfd_factor_tuning": (_I, [_I, _I, _I]),
    "gfdfd_factors_free": (_V, [C.POINTER(_P)]),
    "gfdfd_factors_singular": (_I, [_P, _P]),
    "gfdfd_solve": (_I, [_P, _I, _I, _I, _P, _P, _P]),
    "gfdfd_solve_refine": (_I, [_P, _I, _I, _I, _P, _P, _D, _I, _IP, _DP, _P]),
    "gfdfd_solve_noalloc": (_I, [_P, _I, _I, _I, _P, _P, _P]),
    "gfdfd_solve_noalloc_limit": (C.c_longlong, []),
    "gfdfd_solve_batch": (_I, [_P, _I, _P, _I, _I, _P, _P, _P]),
    "gfdfd_solve_refine_batch": (_I, [_P, _I, _P, _I, _I, _P, _P, _D, _I, _P, _P, _P]),
    "gfdfd_residual": (_I, [_P, _P, _I, _I, _P, _P, _P, _P]),
    "gfdfd_spike_panel": (_I, [_I, _I, _P, _D, _D, _P, _P]),
    "gfdfd_sample": (_I, [_I, _I, _I, _P, _P, _P, _P]),
    "gfdfd_adjoint_sources": (_I, [_I, _I, _I, _P, _P, _P, _F, _P, _P, _P]),
    "gfdfd_residual_sources": (_I, [_I, _I, _I, _P, _P, _P, _F, _I, _I, _I, _P, _P, _P, _P]),
    "gfdfd_grad_ac": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "gfdfd_grad_ac_stf": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gfdfd_grad_ac_taper": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gfdfd_wavelet_amp": (None, [_I, _F, _F, _P, _I, _I, _P]),
    "gfdfd_shot_taper_exponent": (_F, []),
    "gfdfd_hess_shin_ac": (_I, [_I, _I, _P, _P, _P, _P, _P, _P]),
    "gfdfd_hess_shin_te": (_I, [_I, _I, _P, _P, _P, _P, _P, _P]),
    "gfdfd_hess_gn_ac": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "gfdfd_hess_gn_te": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "gfdfd_apply_hess": (_I, [_I, _F, _P, _P, _P]),
    "gfdfd_model_update": (_I, [_I, _F, _P, _P, _F, _F, _P, _P]),
    "gfdfd_model_update_norm": (_I, [_I, _F, _P, _P, _F, _F, _F, _P, _P]),
    "gfdfd_precond": (_I, [_I, _I, _F, _F, _F, _F, _F, _I, _I, _P, _P, _P]),
    "gfdfd_taper_file": (_I, [_I, _I, _I, _I, _P, _P, _P, _P]),
    "gfdfd_grad_frame": (_I, [_I, _I, _I, _I, _P, _P]),
    "gfdfd_tikhonov_te": (_I, [_I, _I, _F, _F, _F, _F, _F, _F, _F, _P, _P, _P, _P, _P, _P]),
    "gfdfd_grad_te": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "gfdfd_grad_te_stf": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gfdfd_grad_te_taper": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gfdfd_profile": (_I, [_I]),
    "gfdfd_device_memory": (_I, [_LLP, _LLP, _I]),
    "gfdfd_release_workspace": (_I, []),
    "gfdfd_profile_select": (_I, [C.c_char_p]),
    "gfdfd_profile_reset": (_V, []),
    "gfdfd_profile_read": (_I, [C.c_char_p, _DP, _LLP, _DP, _DP]),
    "umfpack_zi_triplet_to_col": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "umfpack_zi_symbolic": (_I, [_I, _I, _P, _P, _P, _P, C.POINTER(_P), _P, _P]),
    "umfpack_zi_numeric": (_I, [_P, _P, _P, _P, _P, C.POINTER(_P), _P, _P]),
    "umfpack_zi_solve": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "umfpack_zi_free_symbolic": (_V, [C.POINTER(_P)]),
    "umfpack_zi_free_numeric": (_V, [C.POINTER(_P)]),
}


def lib():
    """The loaded libgfdfd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GfdfdError(f"{LIB_PATH} is missing: build it with `make -C germaine_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, what=""):
    """Raise GfdfdError for a negative status; return the status otherwise."""
    if status < 0:
        msg = lib().gfdfd_last_error().decode(errors="replace")
        raise GfdfdError(f"{what}: status {status}: {msg}")
    return status


def build(jobs=8):
    """Compile libgfdfd.so in-tree for gfx950 (always recompiles what changed)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", os.path.join(_HERE, "csrc"), f"-j{jobs}"])
