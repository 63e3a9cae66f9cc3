"""ctypes binding of libefe_gpu.so (include/efe_gpu.h).

The library is built in-tree (edelweissfe_amd/lib/libefe_gpu.so) by
``edelweissfe_amd.build.build()``.  There is no fallback: if the library is
missing or cannot be loaded, every entry point raises.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libefe_gpu.so")

EFE_OK = 0
EFE_CUTBACK = 1
EFE_NAN = 2
EFE_NOT_CONVERGED = 3
EFE_ERR_ARG, EFE_ERR_HIP, EFE_ERR_STATE = -1, -2, -3

EFE_C3D8, EFE_C3D20, EFE_C3D20R = 0, 1, 2
EFE_LINEAR_ELASTIC, EFE_VONMISES, EFE_DRUCKER_PRAGER = 0, 1, 2
EFE_FP64, EFE_FP32, EFE_FP32_MATFREE = 0, 1, 2
EFE_SOLVER_CG, EFE_SOLVER_GMRES, EFE_SOLVER_BICGSTAB = 0, 1, 2
EFE_PRECOND_BLOCK_JACOBI, EFE_PRECOND_ILU0 = 0, 1

ELEMENT_TYPES = {"C3D8": EFE_C3D8, "C3D20": EFE_C3D20, "C3D20R": EFE_C3D20R}
MATERIALS = {"linearelastic": EFE_LINEAR_ELASTIC, "vonmises": EFE_VONMISES, "druckerprager": EFE_DRUCKER_PRAGER}


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("n_nodes", ctypes.c_int32),
        ("n_elements", ctypes.c_int32),
        ("element_type", ctypes.c_int32),
        ("material", ctypes.c_int32),
        ("coords", ctypes.c_void_p),
        ("conn", ctypes.c_void_p),
        ("n_props", ctypes.c_int32),
        ("props_per_element", ctypes.c_int32),
        ("props", ctypes.c_void_p),
        ("specimen_of_node", ctypes.c_void_p),
        ("n_specimens", ctypes.c_int32),
        ("stream", ctypes.c_void_p),
        ("material_of_element", ctypes.c_void_p),
    ]


class SolveOpts(ctypes.Structure):
    _fields_ = [
        ("solver", ctypes.c_int32),
        ("precond", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("max_iter", ctypes.c_int32),
        ("restart", ctypes.c_int32),
        ("check_every", ctypes.c_int32),
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
    ]


class NewtonOpts(ctypes.Structure):
    """efe_newton_opts"""
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("krylov_max_iter", ctypes.c_int32),
        ("check_every", ctypes.c_int32),
        ("mixed_first", ctypes.c_int32),
        ("alt_after", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("rtol", ctypes.c_double),
        ("rtol_first", ctypes.c_double),
        ("atol_factor", ctypes.c_double),
        ("flux_tol", ctypes.c_double),
        ("flux_tol_alt", ctypes.c_double),
        ("correction_tol", ctypes.c_double),
        ("flux_floor", ctypes.c_double),
        ("avg_floor", ctypes.c_double),
    ]


EFE_NEWTON_ITERATING, EFE_NEWTON_CONVERGED, EFE_NEWTON_FAILED, EFE_NEWTON_IDLE = 0, 1, 2, 3
EFE_FAIL_NONE, EFE_FAIL_CUTBACK, EFE_FAIL_NAN, EFE_FAIL_ITERATIONS, EFE_FAIL_KRYLOV = 0, 1, 2, 3, 4


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("residual", ctypes.c_double),
        ("rhs_norm", ctypes.c_double),
    ]


# name -> (restype, argtypes), exactly the declarations of include/efe_gpu.h
_P = ctypes.c_void_p
_I = ctypes.c_int32
_D = ctypes.c_double
SIGNATURES = {
    "efe_last_error": (ctypes.c_char_p, []),
    "efe_ctx_create": (_I, [ctypes.POINTER(MeshDesc), ctypes.POINTER(_P)]),
    "efe_ctx_destroy": (None, [_P]),
    "efe_get_sizes": (_I, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "efe_get_csr_pattern": (_I, [_P, _P, _P]),
    "efe_set_dirichlet": (_I, [_P, _P, _I]),
    "efe_compute_elements": (_I, [_P, _P, _P, _P, _D, _P, _P, _P, _P, _I, ctypes.POINTER(_D)]),
    "efe_compute_residual": (_I, [_P, _P, _P, _P, _D, _P, _P, _I, ctypes.POINTER(_D)]),
    "efe_body_force": (_I, [_P, _P, _I, _P, _P]),
    "efe_surface_traction": (_I, [_P, _P, _P, _I, _P, _P]),
    "efe_collect_results": (_I, [_P, _I, _P, _I, _P, _I, _I, _I, _P]),
    "efe_get_element_matrices": (_I, [_P, _P, _I, _P, _P]),
    "efe_set_geostatic": (_I, [_P, ctypes.POINTER(_D), _P, _I]),
    "efe_export_csr_values": (_I, [_P, _P]),
    "efe_spmv": (_I, [_P, _P, _P]),
    "efe_linsolve": (_I, [_P, _P, _P, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    "efe_convergence_norms": (_I, [_P, _P, _P, _P, ctypes.POINTER(_D)]),
    "efe_convergence_norms_batched": (_I, [_P, _P, _P, _P, _P]),
    "efe_linsolve_batched": (_I, [_P, _P, _P, ctypes.POINTER(SolveOpts), _P, _P, _P, _P, ctypes.POINTER(SolveInfo)]),
    "efe_accept_state_batched": (_I, [_P, _P]),
    "efe_newton_round_batched": (_I, [_P] + [_P] * 11 + [ctypes.POINTER(NewtonOpts)] + [_P] * 4),
    "efe_get_specimen_count": (_I, [_P]),
    "efe_get_specimen_status": (_I, [_P, _P]),
    "efe_accept_state": (_I, [_P]),
    "efe_get_qp_state": (_I, [_P, _I, _P]),
    "efe_set_qp_state": (_I, [_P, _P]),
    "efe_set_active_specimens": (_I, [_P, _P]),
    "efe_set_launch": (_I, [_P, _I, _I]),
    "efe_get_phase_times": (_I, [_P, ctypes.POINTER(_D)]),
    "efe_element_compute": (_I, [_P, _P, _I, _P, _P, _P, _P, _P, _P, ctypes.POINTER(_D), _D, ctypes.POINTER(_D)]),
    "efe_csr_linsolve": (_I, [_I, _P, _P, _P, _P, _P, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    "efe_csr_linsolve_device": (_I, [_I, _P, _P, _P, _P, _P, _P, ctypes.POINTER(SolveOpts),
                                     ctypes.POINTER(SolveInfo)]),
    "efe_csr_linsolve_device_noalloc": (_I, [_I, _P, _P, _P, _P, _P, _P, ctypes.POINTER(SolveOpts),
                                              ctypes.POINTER(SolveInfo)]),
}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load():
    """Load libefe_gpu.so; raises LibraryMissing when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            "libefe_gpu.so not found at {}; build it with `python -m edelweissfe_amd.build`".format(LIB_PATH))
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class EfeError(RuntimeError):
    pass


def check(rc, what):
    if rc < 0:
        msg = load().efe_last_error()
        raise EfeError("{} failed ({}): {}".format(what, rc, msg.decode() if msg else ""))
    return rc
