"""ctypes binding of libbielcc_hip.so (the C ABI in include/bielcc_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback behind this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbielcc_hip.so")

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


class NumParam(C.Structure):
    """bielcc_num_param: NumParam (electrodynamics/num_param.h:28-55)."""
    _fields_ = [("eps", C.c_double), ("rs", C.c_double), ("rs_seg", C.c_double),
                ("analytic_calc", C.c_double), ("n_start", C.c_int), ("n_start_seg", C.c_int),
                ("p_max", C.c_int), ("p_max_seg", C.c_int)]


class KernelParam(C.Structure):
    """bielcc_kernel_param: KernelParam (integrators/kernel_param.h:25-57)."""
    _fields_ = [("smoothR", C.c_double), ("smoothRMin", C.c_double), ("k_re", C.c_double),
                ("k_im", C.c_double), ("vec_dbl", C.c_double * 3), ("n", C.c_double * 3),
                ("current", C.c_double * 6)]


class IntegralParam(C.Structure):
    """bielcc_integral_param: IntegralParam (integrators/integral_param.h:16-45)."""
    _fields_ = [("idim", C.c_int), ("n_start", C.c_int), ("p_max", C.c_int), ("eps", C.c_double)]


class AssembleItem(C.Structure):
    """bielcc_assemble_item."""
    _fields_ = [("mesh_id", C.c_int), ("k_re", C.c_double), ("k_im", C.c_double), ("Z_d", C.c_void_p),
                ("ld", C.c_int64)]


class FieldItem(C.Structure):
    """bielcc_field_item."""
    _fields_ = [("mesh_id", C.c_int), ("k_re", C.c_double), ("k_im", C.c_double), ("nrhs", C.c_int),
                ("jE_d", C.c_void_p), ("inc_d", C.c_void_p), ("multE_re", C.c_double), ("multE_im", C.c_double),
                ("M", C.c_int), ("pts_d", C.c_void_p), ("E_d", C.c_void_p), ("H_d", C.c_void_p)]


# symbol -> (restype, argtypes)
_SIGS = {
    "bielcc_version": (C.c_int, []),
    "bielcc_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "bielcc_ctx_destroy": (C.c_int, [C.c_void_p]),
    "bielcc_sync": (C.c_int, [C.c_void_p]),
    "bielcc_stream": (C.c_void_p, [C.c_void_p]),
    "bielcc_last_error": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "bielcc_malloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "bielcc_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bielcc_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "bielcc_mesh_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _DP, _IP]),
    "bielcc_mesh_destroy": (C.c_int, [C.c_void_p, C.c_int]),
    "bielcc_mesh_tables": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "bielcc_mesh_grid_step": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "bielcc_mesh_size": (C.c_int, [C.c_void_p, C.c_int, _IP, _IP]),
    "bielcc_op_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.POINTER(NumParam), C.c_double, C.c_double, C.c_void_p]),
    "bielcc_kernel_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                      C.POINTER(KernelParam), C.c_void_p]),
    "bielcc_integral_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                        C.POINTER(KernelParam), C.POINTER(IntegralParam), C.c_void_p, C.c_void_p]),
    "bielcc_k0_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.POINTER(KernelParam), C.POINTER(IntegralParam), C.c_double,
                                  C.c_void_p, C.c_void_p]),
    "bielcc_analytic_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                        C.POINTER(IntegralParam), C.c_double, C.c_double, C.c_double, C.c_void_p]),
    "bielcc_plane_wave_e": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, _DP, _DP, C.c_void_p]),
    "bielcc_plane_wave_h": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, _DP, _DP, C.c_double, C.c_double,
                                      C.c_double, C.c_void_p]),
    "bielcc_field_add_incident_h": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, _DP, C.c_double,
                                              C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p]),
    "bielcc_integral_y_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                          C.POINTER(KernelParam), _DP, C.POINTER(IntegralParam), C.c_void_p,
                                          C.c_void_p]),
    "bielcc_assemble": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(NumParam), C.c_double, C.c_double,
                                  C.c_int, C.c_int, C.c_void_p, C.c_int64]),
    "bielcc_rhs_planewave": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _DP, _DP, C.c_void_p, C.c_int64]),
    "bielcc_lu_factor": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, _IP]),
    "bielcc_lu_solve": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int, C.c_void_p,
                                  C.c_int64]),
    "bielcc_currents_by_basis": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p]),
    "bielcc_field": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(NumParam), C.c_double, C.c_double, C.c_void_p,
                               C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, _DP, _DP, C.c_int,
                               C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bielcc_field_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(NumParam), C.c_double, C.c_double,
                                       C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, _DP,
                                       _DP, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bielcc_get_field_e": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_double, _DP, _DP,
                                     C.POINTER(NumParam), C.c_double, C.c_double, C.c_int, C.c_void_p,
                                     C.c_void_p]),
    "bielcc_field_batch": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(NumParam), C.c_double, C.c_double, C.c_int,
                                     C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, _DP,
                                     C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bielcc_assemble_ex": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(NumParam), C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64]),
    "bielcc_lu32_factor": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                     C.c_void_p, _IP]),
    "bielcc_lu_factor_rbt": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                       C.c_void_p, C.c_void_p, C.c_uint64, _IP]),
    "bielcc_rbt_solve": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                   C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_double, C.c_int, _IP, _DP]),
    "bielcc_lu32_factor_rbt": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                         C.c_void_p, C.c_void_p, C.c_uint64, _IP]),
    "bielcc_refine_solve_rbt": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                          C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_double,
                                          C.c_int, _IP, _DP]),
    "bielcc_refine_solve": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                      C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_double, C.c_int, _IP, _DP]),
    "bielcc_residual": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                   C.c_int64, C.c_int, C.c_int, C.c_void_p, C.c_int64]),
    "bielcc_far_pattern": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(NumParam), C.c_double, C.c_double, C.c_int,
                                     C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bielcc_plan_assemble": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(AssembleItem), C.c_int,
                                       C.POINTER(NumParam), C.c_int, C.POINTER(C.c_void_p)]),
    "bielcc_plan_field": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(FieldItem), C.POINTER(NumParam), C.c_int,
                                    C.c_int, C.POINTER(C.c_void_p)]),
    "bielcc_plan_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bielcc_plan_capture": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bielcc_plan_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bielcc_plan_info": (C.c_int, [C.c_void_p, C.c_void_p, C.c_char_p, _DP]),
    "bielcc_aver_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                                    C.c_void_p]),
    "bielcc_rwg_create": (C.c_int, [C.c_int, C.c_int, _DP, C.POINTER(C.c_void_p)]),
    "bielcc_rwg_destroy": (C.c_int, [C.c_void_p]),
    "bielcc_rwg_sizes": (C.c_int, [C.c_void_p, _IP, _IP, _IP]),
    "bielcc_rwg_get": (C.c_int, [C.c_void_p, _DP, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _DP, _DP, _IP, _IP]),
    "bielcc_rwg_write_vtk": (C.c_int, [C.c_void_p, _DP, C.c_char_p]),
    "bielcc_set_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "bielcc_get_tuning": (C.c_int, [C.c_void_p, C.c_char_p, _IP]),
    "bielcc_timing_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "bielcc_timing_reset": (C.c_int, [C.c_void_p]),
    "bielcc_timing": (C.c_int, [C.c_void_p, C.c_char_p, _DP, _IP]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m bielcc_amd.build` "
                           "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols() -> list[str]:
    return list(_SIGS)
