"""ctypes prototypes of include/hpcrack.h (the C-ABI of libhpcrack_amd.so).

The library is the product path: there is no fallback.  If it is not built,
loading raises immediately.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# HPCRACK_AMD_LIB selects another build of the same library (e.g. one
# compiled with different tuning macros)
LIB_PATH = os.environ.get("HPCRACK_AMD_LIB") or os.path.join(HERE, "libhpcrack_amd.so")

OK = 0
ERR_ARG, ERR_DEVICE, ERR_MEMORY, ERR_IO, ERR_ZERO_DIAGONAL, ERR_STATE = -1, -2, -3, -4, -5, -6

# enum hpcr_field
(U, FU, RU, UOLD, P, W, D, FD, RD, DOLD, H, K, G, GC, KE, GE, FLAG, DPREV, PD, WD) = range(20)
VECTOR_FIELDS = {U, FU, RU, UOLD, P, W}
# enum hpcr_kernel
K_MECH_PCG_A, K_MECH_PCG_B, K_MECH_RESID, K_PF_PCG_A, K_PF_PCG_B = range(5)
HIST_MAX = 64


class MG(ctypes.Structure):
    _fields_ = [("nu0", ctypes.c_int), ("nu1", ctypes.c_int), ("nu2", ctypes.c_int),
                ("gamma", ctypes.c_int), ("ncy", ctypes.c_int)]


class Geometry(ctypes.Structure):
    _fields_ = [("maxlevel", ctypes.c_int), ("element_nb", ctypes.c_int * 3),
                ("xa", ctypes.c_double), ("xb", ctypes.c_double), ("ya", ctypes.c_double),
                ("yb", ctypes.c_double), ("za", ctypes.c_double), ("zb", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [("small_k", ctypes.c_double), ("lc", ctypes.c_double), ("tol_u", ctypes.c_double),
                ("tol_d_rel", ctypes.c_double), ("tol_d_abs", ctypes.c_double), ("mg_u", MG), ("mg_d", MG)]


class Stats(ctypes.Structure):
    _fields_ = [("cycles", ctypes.c_int), ("smoother_iters", ctypes.c_longlong),
                ("r_norm_initial", ctypes.c_double), ("r_norm_final", ctypes.c_double),
                ("history", ctypes.c_double * HIST_MAX)]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double

# name -> (restype, argtypes)
class MixedParams(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int), ("tol", ctypes.c_double), ("cheb_degree", ctypes.c_int),
                ("power_iters", ctypes.c_int), ("cheb_ratio", ctypes.c_double), ("cheb_upper", ctypes.c_double),
                ("replace_every", ctypes.c_int), ("precond_fp64", ctypes.c_int),
                ("element_mfma", ctypes.c_int), ("coarse_iters", ctypes.c_int)]


class FuserStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_longlong), ("groups", ctypes.c_longlong), ("cycles", ctypes.c_longlong),
                ("launches", ctypes.c_longlong), ("problem_launches", ctypes.c_longlong),
                ("max_problems", ctypes.c_int), ("graph_captures", ctypes.c_longlong),
                ("graph_replays", ctypes.c_longlong)]


PROTOTYPES = {
    "hpcr_device_count": (_i, [_ip]),
    "hpcr_ctx_create": (_i, [ctypes.POINTER(Geometry), ctypes.POINTER(Params), _i, ctypes.POINTER(_vp)]),
    "hpcr_ctx_destroy": (_i, [_vp]),
    "hpcr_last_error": (ctypes.c_char_p, [_vp]),
    "hpcr_level_dims": (_i, [_vp, _i, _ip]),
    "hpcr_level_h": (_i, [_vp, _i, _dp]),
    "hpcr_device_bytes": (ctypes.c_longlong, [_vp]),
    "hpcr_synchronize": (_i, [_vp]),
    "hpcr_set_field": (_i, [_vp, _i, _i, _i, _dp]),
    "hpcr_get_field": (_i, [_vp, _i, _i, _i, _dp]),
    "hpcr_fill_field": (_i, [_vp, _i, _i, _i, _d]),
    "hpcr_coarse_vr": (_i, [_vp, _i]),
    "hpcr_u_entire": (_i, [_vp, _d, _i, _dp, ctypes.POINTER(Stats)]),
    "hpcr_phase_field": (_i, [_vp, _i, _dp, ctypes.POINTER(Stats)]),
    "hpcr_u_entire_mixed": (_i, [_vp, _d, _i, _dp, ctypes.POINTER(MixedParams), ctypes.POINTER(Stats)]),
    "hpcr_mixed_resolve": (_i, [_vp, ctypes.POINTER(MixedParams), _dp]),
    "hpcr_estimate_bytes": (_i, [ctypes.POINTER(Geometry), _i, ctypes.POINTER(ctypes.c_longlong)]),
    "hpcr_set_graphs": (_i, [_vp, _i]),
    "hpcr_set_reduction_fence": (_i, [_vp, _i]),
    "hpcr_fuser_create": (_i, [_i, _i, ctypes.POINTER(_vp)]),
    "hpcr_fuser_destroy": (_i, [_vp]),
    "hpcr_fuser_set_graphs": (_i, [_vp, _i]),
    "hpcr_fuser_get_stats": (_i, [_vp, ctypes.POINTER(FuserStats)]),
    "hpcr_set_fuser": (_i, [_vp, _vp]),
    "hpcr_strain_history": (_i, [_vp, _d]),
    "hpcr_voigt_all_level": (_i, [_vp, _i]),
    "hpcr_vr_all_level": (_i, [_vp, _i]),
    "hpcr_fd_h": (_i, [_vp, _i]),
    "hpcr_U_e": (_i, [_vp, _dp]),
    "hpcr_externe_force_sum": (_i, [_vp, _dp]),
    "hpcr_delta_d_max": (_i, [_vp, _dp]),
    "hpcr_d_max": (_i, [_vp, _dp]),
    "hpcr_save_d_old": (_i, [_vp]),
    "hpcr_internal_force": (_i, [_vp, _i, _i, _i]),
    "hpcr_internal_force_f32": (_i, [_vp, _i, _i, _i]),
    "hpcr_prem_u": (_i, [_vp, _i, _i]),
    "hpcr_fld": (_i, [_vp, _i, _i, _i]),
    "hpcr_precondition_m": (_i, [_vp, _i, _i]),
    "hpcr_mp_choice": (_i, [_vp]),
    "hpcr_mp_all_element": (_i, [_vp]),
    "hpcr_mech_setup": (_i, [_vp]),
    "hpcr_pcg_u": (_i, [_vp, _i, _i, _d, _i, _dp]),
    "hpcr_pcg_d": (_i, [_vp, _i, _i, _dp]),
    "hpcr_cycle_u": (_i, [_vp, _i, _d, _i, _dp]),
    "hpcr_cycle_d": (_i, [_vp, _i, _dp]),
    "hpcr_injection": (_i, [_vp, _i, _i, _i]),
    "hpcr_restrict": (_i, [_vp, _i, _i, _i, _i, _i]),
    "hpcr_prolong": (_i, [_vp, _i, _i, _i, _i, _i]),
    "hpcr_dirichlet_bd": (_i, [_vp, _i, _i, _d]),
    "hpcr_timing_enable": (_i, [_vp, _i]),
    "hpcr_timing_reset": (_i, [_vp]),
    "hpcr_timing_query": (_i, [_vp, _i, _i, _dp, ctypes.POINTER(ctypes.c_longlong)]),
}

_lib = None


def load():
    """Load libhpcrack_amd.so, raising if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m hpcrack_amd.build` "
                               "(or __graft_entry__.build()) first; there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
